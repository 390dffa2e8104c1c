// spf_poisson.hip -- Poisson correction (src/poisson_correction.c:73-454),
// the Chebyshev red-black SOR solver (src/poisson.c:17-130) and the imposed
// electrostatic field (src/poisson_correction.c:456-562).
//
// Red-black SOR: within one half-sweep every updated point reads only points
// of the other colour, so the serial loop order of the reference does not
// affect Psi; the device gets the same bits with any schedule that keeps the
// colour order (here: several half-sweeps per launch on LDS tiles).  The
// convergence test needs sum |residual| after every half-sweep: each tile
// writes one partial per half-sweep and a one-workgroup check kernel adds
// them in tile order (fixed, so deterministic) and sets a done flag that makes
// the queued batches return at once.  The sum's rounding differs from the
// serial sum, so the iteration count can differ from the reference only when
// the residual lands within rounding of the threshold.
#include <cmath>

#include "spf_internal.h"

namespace spf {

namespace {

// ---- temporally blocked red-black SOR ----
// A batch launch performs up to SK consecutive half-sweeps.  Each workgroup
// owns an OZ x OR tile of points and holds it with an SK-point halo (clipped
// at the array edges) in registers: lane l of wavefront w keeps columns
// zlo+2l and zlo+2l+1 of tile rows w*RS .. w*RS+RS-1, with their right-hand
// side and update mask.  A half-sweep updates colour n%2, which in every row
// is one of the lane's two columns; the other-colour neighbours come from the
// lane itself, its neighbour lanes (DPP shifts) and, for the first and last
// rows, the adjacent wavefronts through LDS.  At local step j only points at
// least j+1 away from a halo edge that is not an array edge are updated, so
// after the batch every owned point holds exactly the value the
// one-colour-per-sweep order produces.  Psi is read from one buffer and the
// owned points written to the other.  Each workgroup also sums |residual|
// over its owned points for every step (fixed order) into partials[j][tile];
// sor_check_kernel then adds the tiles in order and applies the stopping rule
// of src/poisson.c:125-126 for each half-sweep in turn.
constexpr int SK = 8;                   // half-sweeps per batch = halo width
constexpr int TZ = 128;                 // tile columns: 64 lanes x 2
constexpr int OZ = TZ - 2 * SK;         // owned columns per tile: 112
constexpr int RS = 6;                   // rows per lane
constexpr int SW = 6;                   // wavefronts per tile
constexpr int TRW = RS * SW;            // tile rows: 36
constexpr int OR = TRW - 2 * SK;        // owned rows per tile: 20
constexpr int SB = SW * 64;             // SOR batch workgroup
constexpr int CB = 1024;                // check kernel: 32 lanes per half-sweep, 32 half-sweeps
static_assert(OZ % 2 == 0 && OR % 2 == 0 && SK % 2 == 0 && RS % 2 == 0,
              "tile origins and lane rows must be even so each row's colour column is fixed");

// fixed-order sum over a workgroup: butterfly within each wavefront, then the
// wavefront sums in order (the same operations every launch)
__device__ double block_sum_det(double v, double *wsum) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) wsum[w] = v;
  __syncthreads();
  double t = 0;
  for (int k = 0; k < nw; k++) t += wsum[k];
  __syncthreads();
  return t;
}

// lane l receives lane l-1's value (wave_shr:1); lane 0 receives 0
__device__ inline double from_left(double v) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// lane l receives lane l+1's value (wave_shl:1); lane 63 receives 0
__device__ inline double from_right(double v) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// x / d, correctly rounded, for |x| in (1e-280, 1e280) and a host-checked
// y = RN(1/d): q0 = RN(x*y) and two remainder corrections
// q <- RN(q + RN(x - d*q)*y) give RN(x/d) (Markstein's correction theorem:
// the first step makes q faithful, the second rounds it correctly; nothing
// underflows in that range).  Callers divide directly outside it.
__device__ inline double div_fast(double x, double d, double y) {
  double q = x * y;
  double r = __builtin_fma(-q, d, x);
  q = __builtin_fma(r, y, q);
  r = __builtin_fma(-q, d, x);
  return __builtin_fma(r, y, q);
}

__device__ inline bool div_fast_ok(double x) {
  double ax = fabs(x);
  return ax > 1e-280 && ax < 1e280;
}

struct SorArgs {
  int nz, nr, ldz, left_dir, right_dir;
  double a, e;
  double y_e, y_axis;    // RN(1/e), RN(1/(e - 2*c[0]))
  int use_y;             // 1: both reciprocals are safe for div_fast
  const double *c, *d;   // per-row stencil coefficients
  const double *omega;   // per-half-sweep omega
  const double *rhs;     // may be null
  const double *Pin;
  double *Pout;
  const int16_t *flags;
  double *partials;      // [SK][ntiles] of this batch
  const int64_t *state;  // [0] done
  int64_t n0;
  int k;                 // half-sweeps in this batch
  int force;             // run even if done is set (the exact re-run after convergence)
  int tiles_z, ntiles;
};

// Per-lane constants of one tile.  Point (i, col) of a lane is bit 2*i+col
// of the masks.  A point is updated at local step j when its update-mask bit
// is set and it lies at least j+1 points from every shrinking halo edge:
// zdepth[col] > j and rdepth[i] > j.
struct SorLane {
  unsigned upd;     // update mask: in the array interior, not metal (src/poisson.c:63-70)
  unsigned own;     // point belongs to the tile's owned region
  int zdepth[2];    // distance of each column from the nearest shrinking z edge
  int rdepth[RS];   // distance of each row from the nearest shrinking r edge (wave-uniform)
  bool l_mirror;    // column 0 is iz == 0: L takes R's value
  bool r_mirror[2]; // column is iz == nz-1: R takes L's value
};

// One half-sweep of colour C on the registers of this lane.  Row i of the
// lane is tile row w*RS + i, whose parity is the parity of i; its colour-C
// point is column (i + C) & 1 of the lane's pair.
template <int C>
__device__ inline double sor_half_sweep(const SorArgs &a, const SorLane &ln, double (&P)[RS][2],
                                        const double (&D)[RS][2], const double2 up_row, const double2 dn_row,
                                        const double (&cr)[RS], const double (&dr)[RS], int j, double omega,
                                        bool axis_wave) {
  const double A = a.a, E = a.e;
  unsigned okb = ln.upd & ((ln.zdepth[0] > j ? 0x555u : 0u) | (ln.zdepth[1] > j ? 0xAAAu : 0u));
#pragma unroll
  for (int i = 0; i < RS; i++)
    if (ln.rdepth[i] <= j) okb &= ~(3u << (2 * i));
  double q[RS], x[RS], res[RS];
  bool slow = false;
#pragma unroll
  for (int i = 0; i < RS; i++) {
    const int col = (i + C) & 1;
    double L, R;
    if (col == 0) {
      R = P[i][1];
      L = from_left(R);
      L = ln.l_mirror ? R : L;
      R = ln.r_mirror[0] ? L : R;
    } else {
      L = P[i][0];
      R = from_right(L);
      R = ln.r_mirror[1] ? L : R;
    }
    const double zs = R + L;
    const double up = i < RS - 1 ? P[i + 1 < RS ? i + 1 : i][col] : (col ? up_row.y : up_row.x);
    const double down = i > 0 ? P[i > 0 ? i - 1 : 0][col] : (col ? dn_row.y : dn_row.x);
    const double c = cr[i];
    double residual, diag, y;
    if (i == 0) {
      // the axis row (src/poisson.c:84-90) has no "below" term, added as
      // -0.0 (an exact no-op), a 4*c "above" weight and diagonal e - 2*c
      const double cu = axis_wave ? 4 * c : c;
      const double dn = axis_wave ? -0.0 : dr[i] * down;
      diag = axis_wave ? E - 2 * c : E;
      y = axis_wave ? a.y_axis : a.y_e;
      residual = A * zs + cu * up + dn + diag * P[i][col] - D[i][col];
    } else {
      diag = E;
      y = a.y_e;
      residual = A * zs + c * up + dr[i] * down + E * P[i][col] - D[i][col];
    }
    res[i] = residual;
    x[i] = omega * residual;
    q[i] = div_fast(x[i], diag, y);
    slow |= !div_fast_ok(x[i]);
  }
  if (slow || !a.use_y) {
#pragma unroll
    for (int i = 0; i < RS; i++) {
      const double diag = (i == 0 && axis_wave) ? E - 2 * cr[0] : E;
      if (!a.use_y || !div_fast_ok(x[i])) q[i] = x[i] / diag;
    }
  }
  double acc = 0;
#pragma unroll
  for (int i = 0; i < RS; i++) {
    const int col = (i + C) & 1;
    const bool ok = (okb >> (2 * i + col)) & 1;
    const double nv = P[i][col] - q[i];
    P[i][col] = ok ? nv : P[i][col];
    acc += (ok && ((ln.own >> (2 * i + col)) & 1)) ? fabs(res[i]) : 0.0;
  }
  return acc;
}

__global__ __launch_bounds__(SB) void sor_batch_kernel(SorArgs a) {
  __shared__ double2 xrow[2][SW][2][64];   // [parity][wave][first,last row][lane]
  __shared__ double wsum[SK][SW];
  const int64_t done = a.force ? 0 : a.state[0];
  const int tile = blockIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int z0 = (tile % a.tiles_z) * OZ, r0 = (tile / a.tiles_z) * OR;
  const int zo1 = min(z0 + OZ, a.nz + 1), ro1 = min(r0 + OR, a.nr + 1);
  const int zlo = max(z0 - SK, 0), rlo = max(r0 - SK, 0);
  const int zhi = min(z0 + OZ + SK, a.nz + 1), rhi = min(r0 + OR + SK, a.nr + 1);
  const int LZ = zhi - zlo, LR = rhi - rlo;
  // load this lane's points, right-hand side, flags and row coefficients;
  // every address is clamped into the array so all loads issue together
  double P[RS][2], D[RS][2], cr[RS], dr[RS];
  int16_t fl[RS][2];
  const int iz_e = zlo + 2 * lane;
  const double *rhs = a.rhs ? a.rhs : a.Pin;
#pragma unroll
  for (int i = 0; i < RS; i++) {
    const int irc = min(rlo + w * RS + i, a.nr);
    cr[i] = a.c[irc];
    dr[i] = a.d[irc];
#pragma unroll
    for (int col = 0; col < 2; col++) {
      const int64_t p = fix(a.ldz, min(iz_e + col, a.nz), irc);
      P[i][col] = a.Pin[p];
      D[i][col] = rhs[p];
      fl[i][col] = a.flags[p];
    }
  }
  if (done) return;
  SorLane ln;
  ln.upd = 0;
  ln.own = 0;
  const bool sz_lo = zlo > 0, sz_hi = zhi < a.nz + 1, sr_lo = rlo > 0, sr_hi = rhi < a.nr + 1;
#pragma unroll
  for (int col = 0; col < 2; col++) {
    const int lz = 2 * lane + col;
    ln.zdepth[col] = min(sz_lo ? lz : SK + 1, sz_hi ? LZ - 1 - lz : SK + 1);
    ln.r_mirror[col] = iz_e + col == a.nz - 1;
  }
  ln.l_mirror = iz_e == 0;
#pragma unroll
  for (int i = 0; i < RS; i++) {
    const int lr = w * RS + i, ir = rlo + lr;
    ln.rdepth[i] = min(sr_lo ? lr : SK + 1, sr_hi ? LR - 1 - lr : SK + 1);
#pragma unroll
    for (int col = 0; col < 2; col++) {
      const int iz = iz_e + col;
      const bool in = lr < LR && 2 * lane + col < LZ;
      const bool in_range =
          (iz >= 1 && iz <= a.nz - 2) || (iz == 0 && a.left_dir) || (iz == a.nz - 1 && a.right_dir);
      const bool upd = in && ir <= a.nr - 2 && iz <= a.nz - 1 && in_range &&
                       !(fl[i][col] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE));
      if (upd) ln.upd |= 1u << (2 * i + col);
      if (iz >= z0 && iz < zo1 && ir >= r0 && ir < ro1) ln.own |= 1u << (2 * i + col);
      D[i][col] = (upd && a.rhs) ? D[i][col] : 0.0;
    }
  }
  const bool axis_wave = rlo == 0 && w == 0;
  double ares[SK];
#pragma unroll
  for (int j = 0; j < SK; j++) ares[j] = 0;
  for (int j = 0; j < a.k; j++) {
    const int par = j & 1;
    // publish the first and last rows, then read the neighbouring wavefronts' rows
    xrow[par][w][0][lane] = make_double2(P[0][0], P[0][1]);
    xrow[par][w][1][lane] = make_double2(P[RS - 1][0], P[RS - 1][1]);
    __syncthreads();
    const double2 up_row = w + 1 < SW ? xrow[par][w + 1][0][lane] : make_double2(0.0, 0.0);
    const double2 dn_row = w > 0 ? xrow[par][w - 1][1][lane] : make_double2(0.0, 0.0);
    const double omega = a.omega[a.n0 + j];
    double acc;
    if (((a.n0 + j) & 1) == 0)
      acc = sor_half_sweep<0>(a, ln, P, D, up_row, dn_row, cr, dr, j, omega, axis_wave);
    else
      acc = sor_half_sweep<1>(a, ln, P, D, up_row, dn_row, cr, dr, j, omega, axis_wave);
#pragma unroll
    for (int jj = 0; jj < SK; jj++)
      if (jj == j) ares[jj] = acc;
  }
  // fixed-order sums of the owned residuals, one per half-sweep: butterflies
  // within each wavefront, then the wavefronts in order
#pragma unroll
  for (int j = 0; j < SK; j++) {
    double v = ares[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) wsum[j][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < a.k) {
    double tot = 0;
    for (int k = 0; k < SW; k++) tot += wsum[threadIdx.x][k];
    a.partials[(int64_t)threadIdx.x * a.ntiles + tile] = tot;
  }
#pragma unroll
  for (int i = 0; i < RS; i++) {
#pragma unroll
    for (int col = 0; col < 2; col++)
      if ((ln.own >> (2 * i + col)) & 1) a.Pout[fix(a.ldz, iz_e + col, rlo + w * RS + i)] = P[i][col];
  }
}

struct CheckArgs {
  const double *partials;
  int ntiles, k;
  int64_t n0;
  double accuracy;
  int64_t *state;   // [0] done, [1] iterations
  double *sums;     // [0] drive sum, [1] residual of the last checked half-sweep
};

// applies the stopping rule after each half-sweep of a batch, in order: 32
// lanes per half-sweep add the tile partials (strided, then a fixed butterfly)
__global__ __launch_bounds__(CB) void sor_check_kernel(CheckArgs a) {
  __shared__ double tot[CB / 32];
  if (a.state[0]) return;
  const int j = threadIdx.x >> 5, l = threadIdx.x & 31;
  double s = 0;
  if (j < a.k) {
    const double *row = a.partials + (int64_t)j * a.ntiles;
    // 8 independent loads in flight per lane, then added in order
    for (int t0 = l; t0 < a.ntiles; t0 += 32 * 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        int t = t0 + 32 * u;
        x[u] = t < a.ntiles ? row[t] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) s += x[u];
    }
  }
  for (int off = 16; off > 0; off >>= 1) s += __shfl_xor(s, off, 32);
  if (l == 0) tot[j] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int jj = 0; jj < a.k; jj++) {
    int64_t n = a.n0 + jj;
    double total = tot[jj];
    a.sums[1] = total;
    if (n > 0 && (total < a.accuracy * a.sums[0] || total == 0)) {
      a.state[0] = 1;
      a.state[1] = n;
      return;
    }
  }
}

struct QArgs {
  int nz, nr, ldz, left_neu, right_neu, upper_neu;
  double eps0;
  const double *Ez, *Er;
  const double *Ar_in, *Ar_out, *Az, *volume;
  const int16_t *flags;
  double *Q;
};

// Gauss-law apparent charge per cell, src/poisson_correction.c:141-182
__global__ __launch_bounds__(256) void gauss_q_kernel(QArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)a.nz * a.nr;
  if (i >= n) return;
  int ir = (int)(i / a.nz), iz = (int)(i % a.nz);
  int64_t p = fix(a.ldz, iz, ir);
  double dQ;
  double ez = (a.Ez[fix(a.ldz, iz + 1, ir)] - a.Ez[p]) * a.Az[ir];
  if (ir == 0)
    dQ = (a.Er[fix(a.ldz, iz, ir + 1)] * a.Ar_out[ir] + ez) * a.eps0;
  else
    dQ = (a.Er[fix(a.ldz, iz, ir + 1)] * a.Ar_out[ir] - a.Er[p] * a.Ar_in[ir] + ez) * a.eps0;
  bool surface = (a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (iz == 0 && a.left_neu) ||
                 (iz == a.nz - 1 && a.right_neu) || (ir == a.nr - 1 && a.upper_neu);
  a.Q[p] = surface ? 0.0 : dQ;
}

// Q -> -rho/eps0 driving term, :267-274
__global__ __launch_bounds__(256) void q_rhs_kernel(QArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)a.nz * a.nr;
  if (i >= n) return;
  int ir = (int)(i / a.nz), iz = (int)(i % a.nz);
  int64_t p = fix(a.ldz, iz, ir);
  a.Q[p] = -a.Q[p] / (a.volume[ir] * a.eps0);
}

__global__ __launch_bounds__(256) void zero_psi_kernel(double *Psi, int nz, int nr, int ldz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)nz * nr;
  if (i >= n) return;
  int ir = (int)(i / nz), iz = (int)(i % nz);
  Psi[fix(ldz, iz, ir)] = 0.0;
}

struct EArgs {
  int nz, nr, ldz, left_neu, right_neu, imposed;
  double dz, dr, ff;
  const double *Psi;
  const double *Ez, *Er;  // old (correction) or current imposed arrays
  double *Ezn, *Ern;      // new (correction: other ping-pong buffer; imposed: same arrays)
  const int16_t *flags;
};

// correction: Ez -= -(dPsi)/dz (src/poisson_correction.c:290-308);
// imposed: Ez = -ff*(dPsi)/dz (:524-543); masked points keep their value
__device__ inline double ez_upd(const EArgs &a, int iz, int ir) {
  int64_t p = fix(a.ldz, iz, ir);
  if (a.flags[p] & SPF_FL_EZ_ZERO) return a.Ez[p];
  double d = a.Psi[p] - a.Psi[fix(a.ldz, iz - 1, ir)];
  if (a.imposed) return -a.ff * d / a.dz;
  return a.Ez[p] - -d / a.dz;
}

__device__ inline double er_upd(const EArgs &a, int iz, int ir) {
  int64_t p = fix(a.ldz, iz, ir);
  if (a.flags[p] & SPF_FL_ER_ZERO) return a.Er[p];
  double d = a.Psi[p] - a.Psi[fix(a.ldz, iz, ir - 1)];
  if (a.imposed) return -a.ff * d / a.dr;
  return a.Er[p] - -d / a.dr;
}

__global__ __launch_bounds__(256) void e_from_psi_kernel(EArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(a.nz + 1) * (a.nr + 1);
  if (i >= n) return;
  int ir = (int)(i / (a.nz + 1)), iz = (int)(i % (a.nz + 1));
  int64_t p = fix(a.ldz, iz, ir);
  // Ez: loop ir < nr, iz 1..nz-1; ends mirrored; row nr-1 zeroed
  double ez;
  if (ir == a.nr - 1)
    ez = 0.0;
  else if (ir == a.nr)
    ez = a.Ez[p];
  else if (iz == 0) {
    double v = ez_upd(a, 1, ir);
    ez = a.left_neu ? v : -v;
  } else if (iz == a.nz) {
    double v = ez_upd(a, a.nz - 1, ir);
    ez = a.right_neu ? v : -v;
  } else
    ez = ez_upd(a, iz, ir);
  // Er: loop ir 1..nr, iz < nz; axis mirror; Neumann ends zeroed
  double er;
  if (iz == a.nz)
    er = a.Er[p];
  else if (a.left_neu && iz == 0)
    er = 0.0;
  else if (a.right_neu && iz == a.nz - 1)
    er = 0.0;
  else if (ir == 0)
    er = -er_upd(a, iz, 1);
  else
    er = er_upd(a, iz, ir);
  a.Ezn[p] = ez;
  a.Ern[p] = er;
}

__global__ void sor_reset_kernel(int64_t *state, int32_t *counter, double *sums) {
  state[0] = 0;
  state[1] = -1;
  *counter = 0;
  sums[0] = 0;
  sums[1] = 0;
}

QArgs q_args(Run &R) {
  QArgs q;
  q.nz = R.d.nz;
  q.nr = R.d.nr;
  q.ldz = R.ldz;
  q.left_neu = R.d.left_bc == SPF_NEUMANN;
  q.right_neu = R.d.right_bc == SPF_NEUMANN;
  q.upper_neu = R.d.upper_bc == SPF_NEUMANN;
  q.eps0 = R.d.epsilon_o;
  q.Ez = R.E2[0][R.cur[0]];
  q.Er = R.E2[1][R.cur[1]];
  q.Ar_in = R.volume + (R.d.nr + 1);
  q.Ar_out = R.volume + 2 * (R.d.nr + 1);
  q.Az = R.volume + 3 * (R.d.nr + 1);
  q.volume = R.volume;
  q.flags = R.flags;
  q.Q = R.Q;
  return q;
}

}  // namespace

hipError_t launch_sor_reset(Run &R, hipStream_t s) {
  hipLaunchKernelGGL(sor_reset_kernel, dim3(1), dim3(1), 0, s, R.sor_state, R.red_counter, R.sor_sums);
  return hipGetLastError();
}

int64_t sor_tiles(const Run &R) {
  int64_t tz = (R.d.nz + 1 + OZ - 1) / OZ, tr = (R.d.nr + 1 + OR - 1) / OR;
  return tz * tr;
}

int sor_batch() { return SK; }

hipError_t launch_sor_batch(Run &R, const double *Pin, double *Pout, const double *rhs, int64_t n0, int k,
                            int partial_row, int force, hipStream_t s) {
  SorArgs a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_dir = R.d.left_bc == SPF_DIRICHLET;
  a.right_dir = R.d.right_bc == SPF_DIRICHLET;
  double dz = R.d.dz, dr = R.d.dr;
  a.a = 1 / (dz * dz);
  a.e = -2 * (1 / (dz * dz) + 1 / (dr * dr));
  // reciprocals for div_fast, computed with the device's operands
  const double diag0 = a.e - 2 * R.sor_c_host0;
  a.y_e = 1 / a.e;
  a.y_axis = 1 / diag0;
  auto safe = [](double v) { double m = std::fabs(v); return m > 1e-200 && m < 1e200; };
  a.use_y = safe(a.e) && safe(diag0) && safe(a.y_e) && safe(a.y_axis);
  a.c = R.sor_c;
  a.d = R.sor_d;
  a.omega = R.omega;
  a.rhs = rhs;
  a.Pin = Pin;
  a.Pout = Pout;
  a.flags = R.flags;
  a.partials = R.red + (int64_t)partial_row * sor_tiles(R);
  a.state = R.sor_state;
  a.n0 = n0;
  a.k = k;
  a.force = force;
  a.tiles_z = (R.d.nz + 1 + OZ - 1) / OZ;
  a.ntiles = (int)sor_tiles(R);
  hipLaunchKernelGGL(sor_batch_kernel, dim3((unsigned)a.ntiles), dim3(SB), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sor_check(Run &R, int64_t n0, int k, double accuracy, hipStream_t s) {
  CheckArgs c;
  c.partials = R.red;
  c.ntiles = (int)sor_tiles(R);
  c.k = k;
  c.n0 = n0;
  c.accuracy = accuracy;
  c.state = R.sor_state;
  c.sums = R.sor_sums;
  hipLaunchKernelGGL(sor_check_kernel, dim3(1), dim3(CB), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_gauss_q(Run &R, hipStream_t s) {
  QArgs q = q_args(R);
  int64_t n = (int64_t)q.nz * q.nr;
  hipLaunchKernelGGL(gauss_q_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_q_to_rhs(Run &R, hipStream_t s) {
  QArgs q = q_args(R);
  int64_t n = (int64_t)q.nz * q.nr;
  hipLaunchKernelGGL(q_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_zero_psi(Run &R, hipStream_t s) {
  int64_t n = (int64_t)R.d.nz * R.d.nr;
  hipLaunchKernelGGL(zero_psi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.Psi, R.d.nz, R.d.nr,
                     R.ldz);
  return hipGetLastError();
}

hipError_t launch_field_from_psi(Run &R, int imposed, double field_factor, hipStream_t s) {
  EArgs a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_neu = R.d.left_bc == SPF_NEUMANN;
  a.right_neu = R.d.right_bc == SPF_NEUMANN;
  a.imposed = imposed;
  a.dz = R.d.dz;
  a.dr = R.d.dr;
  a.ff = field_factor;
  a.flags = R.flags;
  if (imposed) {
    a.Psi = R.PsiImp;
    a.Ez = R.EzImp;
    a.Er = R.ErImp;
    a.Ezn = R.EzImp;
    a.Ern = R.ErImp;
  } else {
    a.Psi = R.Psi;
    a.Ez = R.E2[0][R.cur[0]];
    a.Er = R.E2[1][R.cur[1]];
    a.Ezn = R.E2[0][1 - R.cur[0]];
    a.Ern = R.E2[1][1 - R.cur[1]];
  }
  int64_t n = (int64_t)(a.nz + 1) * (a.nr + 1);
  hipLaunchKernelGGL(e_from_psi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !imposed) {
    R.cur[0] = 1 - R.cur[0];
    R.cur[1] = 1 - R.cur[1];
  }
  return e;
}

}  // namespace spf
