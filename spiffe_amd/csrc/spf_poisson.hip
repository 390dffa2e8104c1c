// spf_poisson.hip -- Poisson correction (src/poisson_correction.c:73-454),
// the Chebyshev red-black SOR solver (src/poisson.c:17-130) and the imposed
// electrostatic field (src/poisson_correction.c:456-562).
//
// Red-black SOR: within one half-sweep every updated point reads only points
// of the other colour, so the serial loop order of the reference does not
// affect Psi; the device updates one colour per launch and gets the same
// bits.  The convergence test needs sum |residual| after every half-sweep:
// each block writes a partial sum, the last block to finish adds the partials
// in block order (fixed, so deterministic) and sets a done flag that makes
// the remaining queued half-sweeps return at once.  The host checks the flag
// once per batch of launches.  The sum's rounding differs from the serial
// sum, so the iteration count can differ from the reference only when the
// residual lands within rounding of the threshold.
#include "spf_internal.h"

namespace spf {

namespace {

// ---- temporally blocked red-black SOR ----
// A batch launch performs up to SK consecutive half-sweeps.  Each workgroup
// owns an STZ x STR tile of points, loads it with an SK-point halo into LDS
// (clipped at the array edges), and updates colour n%2 for n = n0 .. n0+k-1
// in LDS.  At local step j only points at least j+1 away from a halo edge
// that is not an array edge are updated, so after the batch every owned
// point holds exactly the value the one-colour-per-sweep order produces.
// Psi is read from one buffer and the owned points written to the other.
// Each workgroup also sums |residual| over its owned points for every step
// (fixed tree order) into partials[j][tile]; sor_check_kernel then adds the
// tiles in order and applies the stopping rule of src/poisson.c:125-126 for
// each half-sweep of the batch in turn.
constexpr int STZ = 128, STR = 16, SK = 8;
constexpr int SLZ = STZ + 2 * SK, SLR = STR + 2 * SK;
constexpr int SB = 1024;   // SOR batch workgroup: 16 wavefronts share one LDS tile
constexpr int CB = 256;    // check kernel

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

struct SorArgs {
  int nz, nr, ldz, left_dir, right_dir;
  double a, e;
  const double *c, *d;   // per-row stencil coefficients
  const double *omega;   // per-half-sweep omega
  const double *rhs;     // may be null
  const double *Pin;
  double *Pout;
  const int16_t *flags;
  double *partials;      // [SK][ntiles]
  const int64_t *state;  // [0] done
  int64_t n0;
  int k;                 // half-sweeps in this batch
  int force;             // run even if done is set (the exact re-run after convergence)
  int tiles_z, ntiles;
};

__global__ __launch_bounds__(SB) void sor_batch_kernel(SorArgs a) {
  __shared__ double P[SLR][SLZ];
  __shared__ double D[SLR][SLZ];
  __shared__ uint8_t M[SLR][SLZ];
  __shared__ double wsum[SB / 64];
  if (!a.force && a.state[0]) return;
  const int tile = blockIdx.x;
  const int z0 = (tile % a.tiles_z) * STZ, r0 = (tile / a.tiles_z) * STR;
  // owned region covers the whole array [0, nz] x [0, nr] so every value is carried over
  const int zo1 = min(z0 + STZ, a.nz + 1), ro1 = min(r0 + STR, a.nr + 1);
  const int zlo = max(z0 - SK, 0), zhi = min(z0 + STZ + SK, a.nz + 1);
  const int rlo = max(r0 - SK, 0), rhi = min(r0 + STR + SK, a.nr + 1);
  const int LZ = zhi - zlo, LR = rhi - rlo;
  for (int t = threadIdx.x; t < LZ * LR; t += SB) {
    int lr = t / LZ, lz = t % LZ;
    int iz = zlo + lz, ir = rlo + lr;
    int64_t p = fix(a.ldz, iz, ir);
    P[lr][lz] = a.Pin[p];
    bool in_range = (iz >= 1 && iz <= a.nz - 2) || (iz == 0 && a.left_dir) || (iz == a.nz - 1 && a.right_dir);
    bool upd = ir <= a.nr - 2 && iz <= a.nz - 1 && in_range && !(a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE));
    M[lr][lz] = upd ? 1 : 0;
    D[lr][lz] = (upd && a.rhs) ? a.rhs[p] : 0.0;
  }
  __syncthreads();
  const double A = a.a, E = a.e;
  for (int j = 0; j < a.k; j++) {
    const int64_t n = a.n0 + j;
    const int color = (int)(n & 1);
    const double omega = a.omega[n];
    // shrink only at halo edges that are not array edges
    const int sz0 = zlo > 0 ? j + 1 : 0, sz1 = zhi < a.nz + 1 ? j + 1 : 0;
    const int sr0 = rlo > 0 ? j + 1 : 0, sr1 = rhi < a.nr + 1 ? j + 1 : 0;
    double ares = 0;
    const int half = LZ / 2 + 1;   // covers both parities of [zlo, zhi)
    for (int t = threadIdx.x; t < half * LR; t += SB) {
      int lr = t / half, kk = t % half;
      int ir = rlo + lr;
      int iz = 2 * (zlo / 2 + kk) + ((color + ir) & 1);   // (iz + ir) % 2 == n % 2
      int lz = iz - zlo;
      if (lz < sz0 || lz >= LZ - sz1 || lr < sr0 || lr >= LR - sr1) continue;
      if (!M[lr][lz]) continue;
      double zs;
      if (iz == 0)
        zs = P[lr][lz + 1] + P[lr][lz + 1];
      else if (iz == a.nz - 1)
        zs = P[lr][lz - 1] + P[lr][lz - 1];
      else
        zs = P[lr][lz + 1] + P[lr][lz - 1];
      double up = P[lr + 1][lz], p0 = P[lr][lz], c = a.c[ir], residual;
      if (ir == 0) {
        residual = A * zs + 4 * c * up + (E - 2 * c) * p0 - D[lr][lz];
        P[lr][lz] = p0 - omega * residual / (E - 2 * c);
      } else {
        residual = A * zs + c * up + a.d[ir] * P[lr - 1][lz] + E * p0 - D[lr][lz];
        P[lr][lz] = p0 - omega * residual / E;
      }
      if (iz >= z0 && iz < zo1 && ir >= r0 && ir < ro1) ares += fabs(residual);
    }
    // fixed-order block sum of the owned residuals of this half-sweep
    double tot = block_sum_det(ares, wsum);
    if (threadIdx.x == 0) a.partials[(int64_t)j * a.ntiles + tile] = tot;
  }
  for (int t = threadIdx.x; t < (zo1 - z0) * (ro1 - r0); t += SB) {
    int w = zo1 - z0;
    int iz = z0 + t % w, ir = r0 + t / w;
    a.Pout[fix(a.ldz, iz, ir)] = P[ir - rlo][iz - zlo];
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
  if (j < a.k)
    for (int t = l; t < a.ntiles; t += 32) s += a.partials[(int64_t)j * a.ntiles + t];
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
  int64_t tz = (R.d.nz + 1 + STZ - 1) / STZ, tr = (R.d.nr + 1 + STR - 1) / STR;
  return tz * tr;
}

int sor_batch() { return SK; }

hipError_t launch_sor_batch(Run &R, const double *Pin, double *Pout, const double *rhs, int64_t n0, int k, int force,
                            hipStream_t s) {
  SorArgs a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_dir = R.d.left_bc == SPF_DIRICHLET;
  a.right_dir = R.d.right_bc == SPF_DIRICHLET;
  double dz = R.d.dz, dr = R.d.dr;
  a.a = 1 / (dz * dz);
  a.e = -2 * (1 / (dz * dz) + 1 / (dr * dr));
  a.c = R.sor_c;
  a.d = R.sor_d;
  a.omega = R.omega;
  a.rhs = rhs;
  a.Pin = Pin;
  a.Pout = Pout;
  a.flags = R.flags;
  a.partials = R.red;
  a.state = R.sor_state;
  a.n0 = n0;
  a.k = k;
  a.force = force;
  a.tiles_z = (R.d.nz + 1 + STZ - 1) / STZ;
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
