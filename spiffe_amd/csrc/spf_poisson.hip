// spf_poisson.hip -- Poisson correction (src/poisson_correction.c:73-454),
// the Chebyshev red-black SOR solver (src/poisson.c:17-130) and the imposed
// electrostatic field (src/poisson_correction.c:456-562).
//
// Red-black SOR: within one half-sweep every updated point reads only points
// of the other colour, so the serial loop order of the reference does not
// affect Psi; the device gets the same bits with any schedule that keeps the
// colour order (here: several half-sweeps per launch on register tiles).  The
// convergence test needs sum |residual| after every half-sweep: each tile
// writes one partial per half-sweep; one extra workgroup of the next batch
// launch (or a check launch after the last batch) adds them in tile order
// (fixed, so deterministic) and sets a done flag that makes the queued
// batches return at once.  The sum's rounding differs from the serial sum,
// so the rule decides on the device only when a rounding bound makes the
// outcome certain, and otherwise leaves that half-sweep to the host, which
// decides it from the serial sums (sor_stop_rule).
#include <cmath>
#include <cstring>

#include "spf_internal.h"

namespace spf {

namespace {

// ---- temporally blocked red-black SOR ----
// A batch performs up to SK consecutive half-sweeps.  Each workgroup owns an
// OZ x OR tile of points and holds it with an SK-point halo (clipped at the
// array edges) in registers: lane l of wavefront w keeps columns zlo+2l and
// zlo+2l+1 of tile rows w*RS .. w*RS+RS-1, with their right-hand side and
// update mask.  A half-sweep updates colour n%2, which in every row is one of
// the lane's two columns; the other-colour neighbours come from the lane
// itself, its neighbour lanes (DPP shifts) and, for the first and last rows,
// the adjacent wavefronts through LDS.  At local step j only points at least
// j+1 away from a halo edge that is not an array edge are updated, so after
// the batch every owned point holds exactly the value the
// one-colour-per-sweep order produces.  Each workgroup also sums |residual|
// over its owned points for every step in a fixed order.
// sor_batch_kernel runs one batch per launch, reading Psi from one buffer
// and writing the owned points to another; its extra workgroup (or
// sor_check_kernel after the last batch) applies the stopping rule.
template <int SK_, int RS_, int SW_>
struct SorCfg {
  static constexpr int SK = SK_;            // half-sweeps per batch = halo width
  static constexpr int RS = RS_;            // rows per lane
  static constexpr int SW = SW_;            // wavefronts per tile
  static constexpr int TZ = 128;            // tile columns: 64 lanes x 2
  static constexpr int OZ = TZ - 2 * SK;    // owned columns per tile
  static constexpr int TRW = RS * SW;       // tile rows
  static constexpr int OR = TRW - 2 * SK;   // owned rows per tile
  static constexpr int NT = SW * 64;        // workgroup size
  static constexpr int SWL = SW * 32;       // residual slots of a workgroup per step (lane pairs)
  static_assert(OZ % 2 == 0 && OR % 2 == 0 && SK % 2 == 0 && RS % 2 == 0,
                "tile origins and lane rows must be even so each row's colour column is fixed");
  static_assert(OR >= SK && OZ >= SK, "a halo must come from the adjacent tiles only");
  static_assert(2 * RS <= 16, "one mask bit per point, two masks in a word");
  static_assert(SK * 32 <= NT, "the block sums take 32 threads per half-sweep");
  // mask bits of the lane's even (column 0) and odd (column 1) points
  static constexpr unsigned COL0 = 0x55555555u & ((1u << (2 * RS)) - 1u);
  static constexpr unsigned COL1 = 0xAAAAAAAAu & ((1u << (2 * RS)) - 1u);
};
// Tile shapes (half-sweeps per batch, rows per lane, wavefronts), fixed at
// build time; each solve takes one of two by the grid points it covers
// (sor_shape_for: the run's, or the sum over the runs of a ragged solve).
// Small solves run 12 half-sweeps per batch on 104 x 24 owned points (more
// tiles to fill the GPU, fewer launches); from SPF_SOR_LARGE_POINTS points up
// they run 8 half-sweeps on 112 x 32 owned points (a smaller halo share of the
// work).  Both have 8 wavefronts, 48 tile rows; the large shape never needs
// more tiles, mask words or partial rows than the small one.  Measured per
// half-sweep on the 1.6-cell gun: at 1051 x 421 lines the small shape is 16%
// faster, at 2101 x 841 the large one 7%, at 4201 x 1681 the large one 13%.
#ifndef SPF_SOR_SHAPE_S
#define SPF_SOR_SHAPE_S 12, 6, 8
#endif
#ifndef SPF_SOR_SHAPE_L
#define SPF_SOR_SHAPE_L 8, 6, 8
#endif
#ifndef SPF_SOR_LARGE_POINTS
#define SPF_SOR_LARGE_POINTS 1200000
#endif
using SorSmall = SorCfg<SPF_SOR_SHAPE_S>;
using SorLarge = SorCfg<SPF_SOR_SHAPE_L>;
// wavefronts per SIMD the batch kernels are compiled for (sets their register budget)
#ifndef SPF_SOR_WAVES
#define SPF_SOR_WAVES 4
#endif
constexpr int CB = 1024;              // check kernel: 32 lanes per half-sweep, 32 half-sweeps
// batches per stopping-rule check: as many as one check launch covers, at most 4
template <class G>
constexpr int check_group_of() {
  return (CB / 32) / G::SK < 4 ? (CB / 32) / G::SK : 4;
}
static_assert(check_group_of<SorSmall>() >= 1 && check_group_of<SorLarge>() >= 1,
              "a check covers at least one batch of half-sweeps");
static_assert(SorLarge::OZ >= SorSmall::OZ && SorLarge::OR >= SorSmall::OR && SorLarge::NT <= SorSmall::NT &&
                  SorLarge::SK <= SorSmall::SK,
              "buffers sized for the small shape must hold the large shape's tiles, masks and partial rows");

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
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// lane l receives lane l+1's value (wave_shl:1); lane 63 receives 0
__device__ inline double from_right(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// x / d, correctly rounded, for a host-checked d < 0 with y = RN(1/d), and
// x = +-0 or |x| in (1e-280, 1e280): q0 = RN(x*y) and two remainder
// corrections q <- RN(q + RN(x - d*q)*y) give RN(x/d) (Markstein's correction
// theorem: the first step makes q faithful, the second rounds it correctly;
// nothing underflows in that range).  For x = +-0, q0 is the correctly signed
// zero, each remainder is +0 and +0*y = -0 (y < 0) leaves q0 unchanged.
// Zero residuals are common (metal, and vacuum the solve has not reached), so
// they stay on the fast path.  The batch kernel checks the set after the fact
// (sor_lane_exact) and redoes a tile's batch with plain division when it was
// left.
__device__ inline double div_fast(double x, double d, double y) {
  double q = x * y;
  double r = __builtin_fma(-q, d, x);
  q = __builtin_fma(r, y, q);
  r = __builtin_fma(-q, d, x);
  return __builtin_fma(r, y, q);
}

// frexp exponent e of x (|x| in [2^(e-1), 2^e); 0 for x = +-0, inf and NaN),
// offset so that e in [-929, 930], i.e. |x| in [2^-930, 2^930), a subset of
// (1e-280, 1e280), maps to [0, kExpSpan] and everything else above it
constexpr unsigned kExpOffset = 929, kExpSpan = 1859;
__device__ inline unsigned div_fast_exp(double x) {
  return (unsigned)(__builtin_amdgcn_frexp_exp(x) + (int)kExpOffset);
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup b runs on XCD b % 8).
// Tile order is row-major, so giving XCD x the contiguous run of tiles
// [start_x, start_x + count_x) puts a band of whole tile rows on each XCD:
// a tile's halo was mostly written by tiles of the same XCD in the previous
// batch and is still in that XCD's L2, as are the band's constant rhs and
// flags.  The map is a bijection of [0, n).
__device__ inline int xcd_band_tile(int b, int n) {
  const int x = b & 7, slot = b >> 3, q = n >> 3, rem = n & 7;
  return x * q + min(x, rem) + slot;
}

struct SorCommon {
  int nz, nr, ldz, left_dir, right_dir;
  double a, e;
  double y_e, y_axis;    // RN(1/e), RN(1/(e - 2*c[0]))
  int use_y;             // 1: both diagonals are negative and their reciprocals safe for div_fast
  const double *c, *d;   // per-row stencil coefficients
  const double *omega;   // per-half-sweep omega
  const double *rhs;     // may be null
  const int16_t *flags;
  const uint32_t *masks;  // [ntiles][NT]: update mask | ownership mask << 16, per lane
  int tiles_z, ntiles;
};

template <class G>
struct SorTile {
  int z0, r0, zo1, ro1, zlo, rlo, LZ, LR;
  bool sz_lo, sz_hi, sr_lo, sr_hi, redge;
  __device__ SorTile(const SorCommon &a, int tile) {
    z0 = (tile % a.tiles_z) * G::OZ;
    r0 = (tile / a.tiles_z) * G::OR;
    zo1 = min(z0 + G::OZ, a.nz + 1);
    ro1 = min(r0 + G::OR, a.nr + 1);
    zlo = max(z0 - G::SK, 0);
    rlo = max(r0 - G::SK, 0);
    const int zhi = min(z0 + G::OZ + G::SK, a.nz + 1), rhi = min(r0 + G::OR + G::SK, a.nr + 1);
    LZ = zhi - zlo;
    LR = rhi - rlo;
    // the halo shrinks only at edges that are not array edges
    sz_lo = zlo > 0;
    sz_hi = zhi < a.nz + 1;
    sr_lo = rlo > 0;
    sr_hi = rhi < a.nr + 1;
    redge = zhi >= a.nz;   // the tile holds iz = nz-1
  }
};

// Per-lane state.  Point (i, col) of a lane is bit 2*i+col of the masks.  A
// point is updated at local step j when its update-mask bit is set and it
// lies at least j+1 points from every shrinking halo edge.
template <class G>
struct SorLane {
  double P[G::RS][2], D[G::RS][2];
  double cr[G::RS], dr[G::RS];   // row coefficients c, d of the lane's rows (uniform over the wavefront)
  unsigned eacc;      // largest div_fast_exp of an updated point's x in the batch
  unsigned upd;       // in the array interior and not metal (src/poisson.c:58-120)
  unsigned own;       // in the tile's owned region
  int zdepth[2];      // distance of each column from the nearest shrinking z edge
  bool r_mirror[2];   // column is iz == nz-1: its right neighbour mirrors the left one
  uint64_t updm[G::RS][2];   // the update mask as wavefront masks (bit = lane), uniform
  uint64_t colown[2]; // lanes whose column 0 / 1 lies in the owned columns
  unsigned ownrows;   // bit i: lane row i lies in the owned rows (uniform)
  bool full;          // every point of the wavefront's rows is updated (uniform)
};

// Column pair of a lane: the lane's two columns iz_e, iz_e+1 (iz_e even) are
// adjacent in a row, and ldz is even, so both come in one 16-byte access
// (flags: one 4-byte access).  iz_e is clamped to ldz-2 so the pair stays in
// the row's storage; columns past nz are never updated or owned.
__device__ inline int64_t sor_pair(const SorCommon &a, int iz_e, int ir) {
  return fix(a.ldz, min(iz_e, a.ldz - 2), min(ir, a.nr));
}

// this lane's points from src; every address is clamped into the array so
// all loads issue together (points outside the tile are never updated)
template <class G>
__device__ inline void sor_load_P(const SorCommon &a, const SorTile<G> &t, int w, int lane, const double *src,
                                  double (&P)[G::RS][2]) {
  const int iz_e = t.zlo + 2 * lane;
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const double2 v = *reinterpret_cast<const double2 *>(src + sor_pair(a, iz_e, t.rlo + w * G::RS + i));
    P[i][0] = v.x;
    P[i][1] = v.y;
  }
}

// Update and ownership masks of this lane's points (bit 2*i+col), from the
// flags and the tile geometry: a point is updated in the array interior and
// not metal (src/poisson.c:58-120), owned in the tile's owned region.  They
// do not change during a solve, so sor_masks_kernel stores them once per
// solve and the batches load one word per lane.
template <class G>
__device__ inline void sor_lane_masks(const SorCommon &a, const SorTile<G> &t, int w, int lane, unsigned &upd_m,
                                      unsigned &own_m) {
  const int iz_e = t.zlo + 2 * lane;
  upd_m = 0;
  own_m = 0;
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const int lr = w * G::RS + i, ir = t.rlo + lr;
    const uint32_t f2 = *reinterpret_cast<const uint32_t *>(a.flags + sor_pair(a, iz_e, ir));
#pragma unroll
    for (int col = 0; col < 2; col++) {
      const int iz = iz_e + col;
      const int16_t fl = (int16_t)(col ? (f2 >> 16) : (f2 & 0xffffu));
      const bool in = lr < t.LR && 2 * lane + col < t.LZ;
      const bool in_range =
          (iz >= 1 && iz <= a.nz - 2) || (iz == 0 && a.left_dir) || (iz == a.nz - 1 && a.right_dir);
      const bool upd = in && ir <= a.nr - 2 && iz <= a.nz - 1 && in_range &&
                       !(fl & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE));
      if (upd) upd_m |= 1u << (2 * i + col);
      if (iz >= t.z0 && iz < t.zo1 && ir >= t.r0 && ir < t.ro1) own_m |= 1u << (2 * i + col);
    }
  }
}

// right-hand side, masks (one stored word per lane, and as wavefront
// masks), ownership and the row coefficients of the wavefront's rows, kept in
// registers for the whole batch; without a right-hand side the loads read src
// (same layout) and are discarded
template <class G>
__device__ inline void sor_setup(const SorCommon &a, const SorTile<G> &t, int tile, int w, int lane, SorLane<G> &ln,
                                 const double *src) {
  const int iz_e = t.zlo + 2 * lane;
  const double *rhs = a.rhs ? a.rhs : src;
  const uint32_t m = a.masks[(int64_t)tile * G::NT + threadIdx.x];
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const double2 v = *reinterpret_cast<const double2 *>(rhs + sor_pair(a, iz_e, t.rlo + w * G::RS + i));
    ln.D[i][0] = a.rhs ? v.x : 0.0;
    ln.D[i][1] = a.rhs ? v.y : 0.0;
  }
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const int irc = min(t.rlo + w * G::RS + i, a.nr);
    ln.cr[i] = a.c[irc];
    ln.dr[i] = a.d[irc];
  }
  ln.eacc = 0;
  ln.upd = m & 0xffffu;
  ln.own = m >> 16;
#pragma unroll
  for (int i = 0; i < G::RS; i++)
#pragma unroll
    for (int col = 0; col < 2; col++)
      if (!((ln.upd >> (2 * i + col)) & 1u)) ln.D[i][col] = 0.0;
#pragma unroll
  for (int i = 0; i < G::RS; i++)
#pragma unroll
    for (int col = 0; col < 2; col++) ln.updm[i][col] = __ballot((ln.upd >> (2 * i + col)) & 1u);
#pragma unroll
  for (int col = 0; col < 2; col++) {
    const int lz = 2 * lane + col;
    ln.zdepth[col] = min(t.sz_lo ? lz : G::SK + 1, t.sz_hi ? t.LZ - 1 - lz : G::SK + 1);
    ln.r_mirror[col] = iz_e + col == a.nz - 1;
    ln.colown[col] = __ballot(iz_e + col >= t.z0 && iz_e + col < t.zo1);
  }
  ln.ownrows = 0;
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const int ir = t.rlo + w * G::RS + i;
    if (ir >= t.r0 && ir < t.ro1) ln.ownrows |= 1u << i;
  }
  ln.full = __ballot(ln.upd != (1u << (2 * G::RS)) - 1u) == 0;
}

// One half-sweep of colour C for a wavefront off the axis, in a tile away
// from the z ends, whose points are all updated (ln.full), whose rows are all
// owned and valid at every step of the batch.  The RS/2 rows that update one
// column share the execution mask (the column's zv), so they are computed in
// one region with no branch between them and their dependent chains
// interleave.  The same operations per point as sor_half_sweep, so the same
// Psi; the residual sum adds the rows of column 0, then those of column 1.
template <class G, int C, bool OWNALL>
__device__ inline double sor_half_sweep_rows(const SorCommon &a, SorLane<G> &ln, uint64_t zv0, uint64_t zv1,
                                             const double2 up_row, const double2 dn_row, double omega) {
  const double A = a.a, E = a.e, y = a.y_e;
  double acc = 0;
#pragma unroll
  for (int col = 0; col < 2; col++) {
    constexpr int NR = G::RS / 2;
    const int first = (col + C) & 1;   // rows first, first+2, ... update this column
    double L[NR], R[NR];
#pragma unroll
    for (int g = 0; g < NR; g++) {
      const int i = first + 2 * g;
      if (col == 0) {
        R[g] = ln.P[i][1];
        L[g] = from_left(R[g]);
      } else {
        L[g] = ln.P[i][0];
        R[g] = from_right(L[g]);
      }
    }
    if (!__builtin_amdgcn_inverse_ballot_w64(col ? zv1 : zv0)) continue;
    double res[NR];
#pragma unroll
    for (int g = 0; g < NR; g++) {
      const int i = first + 2 * g;
      const double zs = R[g] + L[g];
      const double up = i < G::RS - 1 ? ln.P[i + 1 < G::RS ? i + 1 : i][col] : (col ? up_row.y : up_row.x);
      const double down = i > 0 ? ln.P[i > 0 ? i - 1 : 0][col] : (col ? dn_row.y : dn_row.x);
      res[g] = A * zs + ln.cr[i] * up + ln.dr[i] * down + E * ln.P[i][col] - ln.D[i][col];
    }
#pragma unroll
    for (int g = 0; g < NR; g++) {
      const int i = first + 2 * g;
      const double x = omega * res[g];
      ln.P[i][col] = ln.P[i][col] - div_fast(x, E, y);
      ln.eacc = max(ln.eacc, div_fast_exp(x));
      if (OWNALL || ((ln.ownrows >> i) & 1u)) acc += fabs(res[g]);
    }
  }
  return acc;
}

// One half-sweep of colour C on the registers of this lane.  Row i of the
// lane is tile row w*RS + i, whose parity is the parity of i (rlo is even);
// its colour-C point is column (i + C) & 1 of the lane's pair.  rv: the rows
// valid at this step (uniform; the others are skipped); zv0, zv1: the lanes
// whose column 0 / 1 is valid at this step.  A point is computed only on the
// lanes that update it (the update mask and zv, as the execution mask), so
// nothing is selected afterwards.  Returns the sum of |residual| over the
// lane's updated owned points, rows in order.  EDGE: the tile touches the
// array's left (ledge) or right (redge) z end and needs the mirrors; AXM: 0
// the wavefront does not hold the axis row, 1 it does, 2 decided by AX at run
// time.  EXACT: divide directly; otherwise div_fast, with each updated point's
// exponent folded into ln.eacc.
template <class G, int C, bool EDGE, int AXM, bool EXACT>
__device__ inline double sor_half_sweep(const SorCommon &a, SorLane<G> &ln, int lo, int hi, uint64_t zv0, uint64_t zv1,
                                        const double2 up_row, const double2 dn_row, double omega, bool ledge,
                                        bool redge, bool AX) {
  const double A = a.a, E = a.e;
  const bool ax = AXM == 2 ? AX : AXM == 1;
  const bool lane0 = (threadIdx.x & 63) == 0;
  double acc = 0;
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    if (i < lo || i >= hi) continue;
    const int col = (i + C) & 1;
    // the other-colour neighbours in the row; the shifts run on every lane
    double L, R;
    if (col == 0) {
      R = ln.P[i][1];
      L = from_left(R);
      // lane 0's column 0 is updated only as iz == 0 of a tile at the left
      // end, whose left neighbour is the mirror of the right one
      if (EDGE && ledge) L = lane0 ? R : L;
      if (EDGE && redge) R = ln.r_mirror[0] ? L : R;
    } else {
      L = ln.P[i][0];
      R = from_right(L);   // lane 63's column 1 is never updated except as iz == nz-1, mirrored here
      if (EDGE && redge) R = ln.r_mirror[1] ? L : R;
    }
    if (!__builtin_amdgcn_inverse_ballot_w64(ln.updm[i][col] & (col ? zv1 : zv0))) continue;
    const double zs = R + L;
    const double up = i < G::RS - 1 ? ln.P[i + 1 < G::RS ? i + 1 : i][col] : (col ? up_row.y : up_row.x);
    const double c = ln.cr[i];
    double residual, diag, y;
    if (i == 0) {
      // row 0 of the tile may be the axis row, src/poisson.c:84-90: no "below"
      // term (added as -0.0, an exact no-op), a 4*c "above" weight, diagonal e - 2*c
      const double down = col ? dn_row.y : dn_row.x;
      const double cu = ax ? 4 * c : c;
      const double dn = ax ? -0.0 : ln.dr[i] * down;
      diag = ax ? E - 2 * c : E;
      y = ax ? a.y_axis : a.y_e;
      residual = A * zs + cu * up + dn + diag * ln.P[i][col] - ln.D[i][col];
    } else {
      const double down = ln.P[i > 0 ? i - 1 : 0][col];
      diag = E;
      y = a.y_e;
      residual = A * zs + c * up + ln.dr[i] * down + E * ln.P[i][col] - ln.D[i][col];
    }
    const double x = omega * residual;
    const double q = EXACT ? x / diag : div_fast(x, diag, y);
    ln.P[i][col] = ln.P[i][col] - q;
    if (!EXACT) ln.eacc = max(ln.eacc, div_fast_exp(x));
    if (((ln.ownrows >> i) & 1u) && __builtin_amdgcn_inverse_ballot_w64(ln.colown[col])) acc += fabs(residual);
  }
  return acc;
}

// One local step j of a batch, colour C, relaxation factor omega; returns the
// lane's owned residual sum.
template <class G, int C, bool EDGE, int AXM, bool EXACT = false>
__device__ inline double sor_step(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane,
                                  double omega, int j, double2 (*xrow)[G::SW][2][64], bool axis_wave) {
  const int par = j & 1;
  // publish the first and last rows, then read the neighbouring wavefronts' rows
  xrow[par][w][0][lane] = make_double2(ln.P[0][0], ln.P[0][1]);
  xrow[par][w][1][lane] = make_double2(ln.P[G::RS - 1][0], ln.P[G::RS - 1][1]);
  __syncthreads();
  const double2 up_row = w + 1 < G::SW ? xrow[par][w + 1][0][lane] : make_double2(0.0, 0.0);
  const double2 dn_row = w > 0 ? xrow[par][w - 1][1][lane] : make_double2(0.0, 0.0);
  // points valid at this step: the wavefront's share of the valid rows and
  // the column depth
  const int vr0 = t.sr_lo ? j + 1 : 0, vr1 = t.LR - (t.sr_hi ? j + 1 : 0);
  const int lo = min(max(vr0 - w * G::RS, 0), G::RS), hi = min(max(vr1 - w * G::RS, 0), G::RS);
  return sor_half_sweep<G, C, EDGE, AXM, EXACT>(a, ln, lo, hi, __ballot(ln.zdepth[0] > j), __ballot(ln.zdepth[1] > j),
                                                up_row, dn_row, omega, t.zlo == 0, t.redge, axis_wave);
}

// step j of a wavefront that takes sor_half_sweep_rows; the lanes outside
// the owned columns return 0
template <class G, int C, bool OWNALL>
__device__ inline double sor_step_rows(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane,
                                       double omega, int j, double2 (*xrow)[G::SW][2][64]) {
  const int par = j & 1;
  xrow[par][w][0][lane] = make_double2(ln.P[0][0], ln.P[0][1]);
  xrow[par][w][1][lane] = make_double2(ln.P[G::RS - 1][0], ln.P[G::RS - 1][1]);
  __syncthreads();
  const double2 up_row = w + 1 < G::SW ? xrow[par][w + 1][0][lane] : make_double2(0.0, 0.0);
  const double2 dn_row = w > 0 ? xrow[par][w - 1][1][lane] : make_double2(0.0, 0.0);
  const uint64_t zv0 = __ballot(ln.zdepth[0] > j), zv1 = __ballot(ln.zdepth[1] > j);
  double acc;
  if (OWNALL) {
    acc = sor_half_sweep_rows<G, C, true>(a, ln, zv0, zv1, up_row, dn_row, omega);
  } else {
    // rows valid at this step (uniform): all of them, or the per-row loop
    const int vr0 = t.sr_lo ? j + 1 : 0, vr1 = t.LR - (t.sr_hi ? j + 1 : 0);
    const int lo = min(max(vr0 - w * G::RS, 0), G::RS), hi = min(max(vr1 - w * G::RS, 0), G::RS);
    if (lo == 0 && hi == G::RS)
      acc = sor_half_sweep_rows<G, C, false>(a, ln, zv0, zv1, up_row, dn_row, omega);
    else
      acc = sor_half_sweep<G, C, false, 0, false>(a, ln, lo, hi, zv0, zv1, up_row, dn_row, omega, false, false, false);
  }
  return __builtin_amdgcn_inverse_ballot_w64(ln.colown[0]) ? acc : 0.0;
}

// The wavefront's residual sum of step j into wsum[j]: lanes l and l^1 add
// their values (DPP quad_perm [1,0,3,2]) and the even lane stores the pair's
// sum, 32 slots per wavefront; sor_block_sums adds the slots once per batch.
// (A butterfly over the 64 lanes after every step kept each wavefront
// waiting on six dependent LDS round trips per half-sweep.)
template <class G>
__device__ inline void sor_wave_sum(double v, int j, int w, int lane, double (*wsum)[G::SWL]) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
  v += __hiloint2double(hi, lo);
  if ((lane & 1) == 0) wsum[j][w * 32 + (lane >> 1)] = v;
}

// Steps j = 0 .. k-1 of an even n0 in colour pairs (0, 1): one copy of the
// pair loop per tile kind, so interior tiles carry no mirror selects and
// wavefronts off the axis no axis terms.
template <class G, bool EDGE, int AXM>
__device__ inline void sor_pairs(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane, int k,
                                 const double *om, double2 (*xrow)[G::SW][2][64], double (*lres)[G::SWL],
                                 bool axis_wave) {
#pragma unroll 1
  for (int j = 0; j < k; j += 2) {
    sor_wave_sum<G>(sor_step<G, 0, EDGE, AXM>(a, t, ln, w, lane, om[j], j, xrow, axis_wave), j, w, lane, lres);
    if (j + 1 < k)
      sor_wave_sum<G>(sor_step<G, 1, EDGE, AXM>(a, t, ln, w, lane, om[j + 1], j + 1, xrow, axis_wave), j + 1, w,
                      lane, lres);
  }
}

// the pair loop of the wavefronts that take sor_half_sweep_rows
template <class G, bool OWNALL>
__device__ inline void sor_pairs_rows(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane, int k,
                                      const double *om, double2 (*xrow)[G::SW][2][64], double (*lres)[G::SWL]) {
#pragma unroll 1
  for (int j = 0; j < k; j += 2) {
    sor_wave_sum<G>(sor_step_rows<G, 0, OWNALL>(a, t, ln, w, lane, om[j], j, xrow), j, w, lane, lres);
    if (j + 1 < k)
      sor_wave_sum<G>(sor_step_rows<G, 1, OWNALL>(a, t, ln, w, lane, om[j + 1], j + 1, xrow), j + 1, w, lane, lres);
  }
}

// steps one by one with run-time colours: an odd n0 (only from
// spf_poisson_sweeps) and the exact division
template <class G, bool EXACT>
__device__ inline void sor_singles(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane,
                                   int64_t n0, int k, const double *om, double2 (*xrow)[G::SW][2][64],
                                   double (*lres)[G::SWL], bool axis_wave) {
#pragma unroll 1
  for (int j = 0; j < k; j++)
    sor_wave_sum<G>(((n0 + j) & 1)
                        ? sor_step<G, 1, true, 2, EXACT>(a, t, ln, w, lane, om[j], j, xrow, axis_wave)
                        : sor_step<G, 0, true, 2, EXACT>(a, t, ln, w, lane, om[j], j, xrow, axis_wave),
                    j, w, lane, lres);
}

// k <= SK half-sweeps n0 .. n0+k-1 with relaxation factors om[0 .. k-1] (LDS,
// read after each step's barrier); lres[j] gets the workgroup's owned
// residual slots of step j.  The tile kind is uniform over the workgroup,
// the axis flag over the wavefront; every copy runs one barrier per step.
template <class G, bool ROWS>
__device__ inline void sor_sweeps(const SorCommon &a, const SorTile<G> &t, SorLane<G> &ln, int w, int lane,
                                  int64_t n0, int k, const double *om, double2 (*xrow)[G::SW][2][64],
                                  double (*lres)[G::SWL], bool exact) {
  const bool axis_wave = t.rlo == 0 && w == 0;
  const bool edge = t.zlo == 0 || t.redge;
  if (exact) {
    sor_singles<G, true>(a, t, ln, w, lane, n0, k, om, xrow, lres, axis_wave);
  } else if (n0 & 1) {
    sor_singles<G, false>(a, t, ln, w, lane, n0, k, om, xrow, lres, axis_wave);
  } else if (edge) {
    if (axis_wave)
      sor_pairs<G, true, 1>(a, t, ln, w, lane, k, om, xrow, lres, axis_wave);
    else
      sor_pairs<G, true, 0>(a, t, ln, w, lane, k, om, xrow, lres, axis_wave);
  } else {
    // rows valid at every step of the batch: at least SK from each shrinking r edge
    const bool rows_valid = (!t.sr_lo || w * G::RS >= G::SK) && (!t.sr_hi || (w + 1) * G::RS <= t.LR - G::SK);
    if (axis_wave)
      sor_pairs<G, false, 1>(a, t, ln, w, lane, k, om, xrow, lres, axis_wave);
    else if (ROWS && ln.full && rows_valid && ln.ownrows == (1u << G::RS) - 1u)
      sor_pairs_rows<G, true>(a, t, ln, w, lane, k, om, xrow, lres);
    else if (ROWS && ln.full && ln.ownrows != 0)
      sor_pairs_rows<G, false>(a, t, ln, w, lane, k, om, xrow, lres);
    else
      sor_pairs<G, false, 0>(a, t, ln, w, lane, k, om, xrow, lres, axis_wave);
  }
}

// The fast division was valid for every updated point of the batch: each x had
// div_fast_exp <= kExpSpan (|x| inside div_fast's set, or x = +-0, inf or NaN),
// and none was inf or NaN, since div_fast turns those into NaN, which every
// later update of the point keeps.
template <class G>
__device__ inline bool sor_lane_exact(const SorLane<G> &ln) {
  bool ok = ln.eacc <= kExpSpan;
#pragma unroll
  for (int i = 0; i < G::RS; i++)
#pragma unroll
    for (int col = 0; col < 2; col++) ok = ok && __builtin_isfinite(ln.P[i][col]);
  return ok;
}

// the tile's owned residual sum of each step j < k into out[j * stride]:
// butterflies within each wavefront, then the wavefronts in order
template <class G>
__device__ inline void sor_block_sums(int k, double (*wsum)[G::SWL], double *out, int64_t stride) {
  __syncthreads();
  const int j = threadIdx.x >> 5, l = threadIdx.x & 31;
  double tot = 0;
  if (j < k)
    for (int q = 0; q < G::SW; q++) tot += wsum[j][q * 32 + l];
  for (int off = 16; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 32);
  if (j < k && l == 0) out[j * stride] = tot;
}

// owned points to dst: a 16-byte store where both columns are owned
template <class G>
__device__ inline void sor_write_owned(const SorCommon &a, const SorTile<G> &t, const SorLane<G> &ln, int w,
                                       int lane, double *dst) {
  const int iz_e = t.zlo + 2 * lane;
#pragma unroll
  for (int i = 0; i < G::RS; i++) {
    const unsigned o = (ln.own >> (2 * i)) & 3u;
    double *p = dst + fix(a.ldz, iz_e, t.rlo + w * G::RS + i);
    if (o == 3u) {
      *reinterpret_cast<double2 *>(p) = make_double2(ln.P[i][0], ln.P[i][1]);
    } else {
      if (o & 1u) p[0] = ln.P[i][0];
      if (o & 2u) p[1] = ln.P[i][1];
    }
  }
}

struct SorArgs {
  SorCommon m;
  const double *Pin;
  double *Pout;
  double *partials;      // [SK][ntiles] of this batch
  int64_t *state;        // [0] done, [1] iterations
  int64_t n0;
  int k;                 // half-sweeps in this batch
  int force;             // run even if done is set (the exact re-run after convergence)
  // check_k > 0: workgroup ntiles applies the stopping rule to the previous
  // check group, half-sweeps check_n0 .. check_n0+check_k-1 with partial rows
  // check_partials[0 .. check_k-1]
  const double *check_partials;
  double *sums;          // [0] drive sum, [1] residual of the last checked half-sweep
  double accuracy;
  double rel;            // summation-order margin of the stopping rule (sor_stop_rule)
  int64_t check_n0;
  int check_k;
};

// Total of one half-sweep's tile partials by the 32 lanes l of a
// half-wavefront: lane l adds tiles l, l+32, ... (8 loads in flight, added in
// order), then a fixed butterfly; every lane returns the total.  The order
// depends on ntiles alone, so the stopping rule is deterministic.
__device__ inline double sor_partial_total(const double *row, int ntiles, int l) {
  double s = 0;
  for (int t0 = l; t0 < ntiles; t0 += 32 * 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int t = t0 + 32 * u;
      x[u] = t < ntiles ? row[t] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += x[u];
  }
  for (int off = 16; off > 0; off >>= 1) s += __shfl_xor(s, off, 32);
  return s;
}

// The stopping rule (src/poisson.c:125-126) after half-sweeps n0, n0+1, ...
// in order, on their totals tot[0 .. k-1]; one thread.  The reference adds
// |residual| serially (ir outer, iz inner) and the drive sum serially (iz
// outer, ir inner); the device's totals are fixed-order tree sums of the same
// non-negative terms.  Two summation orders of n non-negative terms differ by
// at most rel = 2g/(1-g) times the sum (g = (n-1)u/(1-(n-1)u), u = 2^-53), so
// the comparison is decided here only when it holds for every value the
// serial sums can take: residual total t in [T(1-rel), T(1+rel)], threshold
// RN(accuracy*D) with D in [S(1-rel), S(1+rel)].  Otherwise the solve stops
// with state[2] set, and the host decides that half-sweep from the serial
// sums.  A NaN total (a NaN term, so NaN in any order) never converges; totals
// near the overflow range are left to the host.
__device__ inline void sor_stop_rule(const double *tot, int k, int64_t n0, double accuracy, double rel,
                                     int64_t *state, double *sums) {
  const double drive = sums[0];
  const double u2 = 2.220446049250313e-16;
  const double big = 0x1p1000;
  const double thr_lo = accuracy * (drive * (1 - rel)) * (1 - u2) - 0x1p-1074;
  const double thr_hi = accuracy * (drive * (1 + rel)) * (1 + u2) + 0x1p-1074;
  const bool far = drive < big && accuracy * drive < big;   // NaN drive: every comparison is false
  for (int jj = 0; jj < k; jj++) {
    int64_t n = n0 + jj;
    double total = tot[jj];
    sums[1] = total;
    if (n == 0) continue;
    if (total == 0) {   // every term is zero, in any order
      state[0] = 1;
      state[1] = n;
      return;
    }
    if (total != total || drive != drive) continue;   // NaN stays NaN in any order: no stop
    if (!(total < big) || !far) {
      state[2] = 1;
      state[0] = 1;
      state[1] = n;
      return;
    }
    const double t_hi = total * (1 + rel) * (1 + u2), t_lo = total * (1 - rel) * (1 - u2);
    if (t_hi < thr_lo) {
      state[0] = 1;
      state[1] = n;
      return;
    }
    if (t_lo >= thr_hi) continue;
    state[2] = 1;   // within rounding of the threshold: the host decides from the serial sums
    state[0] = 1;
    state[1] = n;
    return;
  }
}

// all threads of a workgroup of nt threads (a multiple of 32): totals of the
// k half-sweeps into tot, then the stopping rule
__device__ inline void sor_group_check(const double *partials, int ntiles, int k, int64_t n0, double accuracy,
                                       double rel, int64_t *state, double *sums, double *tot, int nt) {
  if (state[0]) return;
  const int g = threadIdx.x >> 5, l = threadIdx.x & 31;
  for (int j0 = 0; j0 < k; j0 += nt / 32) {
    const int j = j0 + g;
    if (j < k) {
      const double tj = sor_partial_total(partials + (int64_t)j * ntiles, ntiles, l);
      if (l == 0) tot[j] = tj;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) sor_stop_rule(tot, k, n0, accuracy, rel, state, sums);
}

// One workgroup of a batch launch: tile `bid` of the batch, or with bid ==
// ntiles the stopping-rule check of the previous check group.  The launch's
// own input and output buffers, half-sweep range, partial rows and check
// range come as arguments, so a ragged launch over several runs can give each
// run its own.
template <class G, bool ROWS = false>
__device__ inline void sor_batch_block(const SorArgs &a, int bid, const double *Pin, double *Pout, int64_t n0, int k,
                                       double *partials, const double *check_partials, int64_t check_n0, int check_k) {
  if (bid == a.m.ntiles) {
    // the extra workgroup: stopping rule of the previous check group, on a
    // CU the tiles leave free (tiles that start after it stops the solve
    // return at once; their output is never used)
    __shared__ double tot[CB / 32];
    sor_group_check(check_partials, a.m.ntiles, check_k, check_n0, a.accuracy, a.rel, a.state, a.sums, tot, G::NT);
    return;
  }
  __shared__ double2 xrow[2][G::SW][2][64];   // [parity][wave][first,last row][lane]
  __shared__ double wsum[G::SK][G::SWL];
  // the done flag, read once for the workgroup (the check workgroup of the
  // same launch may set it at any moment, and the wavefronts must agree), and
  // the batch's relaxation factors, published through LDS with the tile loads
  __shared__ int done;
  __shared__ double om[G::SK];
  if (threadIdx.x == 0) done = a.force ? 0 : (int)a.state[0];
  if (threadIdx.x < G::SK) om[threadIdx.x] = a.m.omega[n0 + min((int)threadIdx.x, k - 1)];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int tile = xcd_band_tile(bid, a.m.ntiles);
  const SorTile<G> t(a.m, tile);
  SorLane<G> ln;
  sor_load_P<G>(a.m, t, w, lane, Pin, ln.P);
  sor_setup<G>(a.m, t, tile, w, lane, ln, Pin);
  // every wavefront reads om[0] before the first step's own barrier: this
  // barrier makes wavefront 0's stores visible to all of them (without it a
  // wavefront that runs ahead reads what an earlier workgroup left there)
  __syncthreads();
  if (done) return;
  sor_sweeps<G, ROWS>(a.m, t, ln, w, lane, n0, k, om, xrow, wsum, !a.m.use_y);
  if (a.m.use_y && __syncthreads_or(!sor_lane_exact<G>(ln))) {
    // rare: some x left div_fast's set; the batch input is intact, so redo
    // the tile's batch with plain division
    sor_load_P<G>(a.m, t, w, lane, Pin, ln.P);
    sor_sweeps<G, ROWS>(a.m, t, ln, w, lane, n0, k, om, xrow, wsum, true);
  }
  sor_block_sums<G>(k, wsum, partials + tile, a.m.ntiles);
  sor_write_owned<G>(a.m, t, ln, w, lane, Pout);
}

// The grouped-row loop (sor_half_sweep_rows) where it was measured faster:
// the large tile shape (3.3% per half-sweep on the 4201 x 1681-line grid, 2.3%
// on the 64-run scan's ragged solves); on the small shape at 1051 x 421 it
// was no faster.
template <class G>
constexpr bool rows_loop() {
  return G::SK == SorLarge::SK && G::RS == SorLarge::RS && G::SW == SorLarge::SW;
}

// Four wavefronts per SIMD (128 VGPRs): two 8-wavefront workgroups share a
// CU, so one hides the other's barrier and DPP latencies.  The compiler keeps
// a few batch constants in scratch to fit; measured 1.37x faster per batch
// on the 4201 x 1681-line grid than at 147 VGPRs and one workgroup per CU.
template <class G>
__global__ __launch_bounds__(G::NT) __attribute__((amdgpu_waves_per_eu(SPF_SOR_WAVES, SPF_SOR_WAVES))) void sor_batch_kernel(
    SorArgs a) {
  sor_batch_block<G, rows_loop<G>()>(a, (int)blockIdx.x, a.Pin, a.Pout, a.n0, a.k, a.partials, a.check_partials, a.check_n0,
                           a.check_k);
}

// A ragged batch launch over several runs' solves in lockstep: workgroups
// [first[r], first[r+1]) belong to run r (each span a multiple of 8, so the
// XCD banding of the tiles is the single-run one); run r's batch reads
// ring[b % 6] and writes ring[(b+1) % 6] of its own buffer ring.  The runs
// share the half-sweep range and the check range; everything else (grid,
// coefficients, right-hand side, partial sums, state) is the run's own.
struct SorMultiRun {
  SorArgs a;               // the run's solve; Pin/Pout/n0/k/partials/check fields unused
  const double *ring[6];
};

struct SorMultiLaunch {
  int64_t n0, b, partial_row, check_row, check_n0;
  int k, check_k, nruns;
};

__device__ inline int ragged_run(const int32_t *first, int n, int b) {
  int lo = 0, hi = n;   // first[lo] <= b < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

template <class G>
__global__ __launch_bounds__(G::NT) __attribute__((amdgpu_waves_per_eu(SPF_SOR_WAVES, SPF_SOR_WAVES))) void sor_multi_kernel(
    const SorMultiRun *__restrict__ T, const int32_t *__restrict__ first, SorMultiLaunch L) {
  const int r = ragged_run(first, L.nruns, (int)blockIdx.x);
  const SorMultiRun &x = T[r];
  const int bid = (int)blockIdx.x - first[r];
  if (bid > x.a.m.ntiles || (bid == x.a.m.ntiles && L.check_k == 0)) return;   // span padding
  const int ntiles = x.a.m.ntiles;
  sor_batch_block<G, rows_loop<G>()>(x.a, bid, x.ring[L.b % 6], const_cast<double *>(x.ring[(L.b + 1) % 6]), L.n0, L.k,
                  x.a.partials + L.partial_row * ntiles, x.a.partials + L.check_row * ntiles, L.check_n0, L.check_k);
}

// the lanes' update and ownership masks of every tile, once per solve
template <class G>
__global__ __launch_bounds__(G::NT) void sor_masks_kernel(SorCommon a, uint32_t *out) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const SorTile<G> t(a, (int)blockIdx.x);
  unsigned upd, own;
  sor_lane_masks<G>(a, t, w, lane, upd, own);
  out[(int64_t)blockIdx.x * G::NT + threadIdx.x] = upd | (own << 16);
}

struct CheckArgs {
  const double *partials;
  int ntiles, k;
  int64_t n0;
  double accuracy, rel;
  int64_t *state;   // [0] done, [1] iterations, [2] left to the host
  double *sums;     // [0] drive sum, [1] residual of the last checked half-sweep
};

// the stopping rule of one check group on its own (the last group of a solve)
__global__ __launch_bounds__(CB) void sor_check_kernel(CheckArgs a) {
  __shared__ double tot[CB / 32];
  sor_group_check(a.partials, a.ntiles, a.k, a.n0, a.accuracy, a.rel, a.state, a.sums, tot, CB);
}

// One half-sweep n of src/poisson.c:58-120 written as the reference writes it,
// in place (a colour-n point reads only points of the other colour and
// itself), with plain division; |residual| of every point of the half-sweep's
// loop goes to dump[ir*nz + iz] (0 where the point is not updated), so a
// serial sum of dump in index order is the reference's abs_total_residual.
// The result equals the batch kernel's half-sweep n bit for bit: the same
// operations in the same order, and the batch's fast division is correctly
// rounded where it is used.
struct ExactSweepArgs {
  SorCommon m;
  double *P;
  double *dump;
  int64_t n;
};

__global__ __launch_bounds__(256) void sor_exact_sweep_kernel(ExactSweepArgs x) {
  const SorCommon &a = x.m;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)a.nz * (a.nr - 1)) return;
  const int ir = (int)(i / a.nz), iz = (int)(i % a.nz);
  double r = 0.0;
  const bool in_range = (iz >= 1 && iz <= a.nz - 2) || (iz == 0 && a.left_dir) || (iz == a.nz - 1 && a.right_dir);
  const int64_t p = fix(a.ldz, iz, ir);
  if (in_range && (ir + iz) % 2 == x.n % 2 && !(a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) {
    double *P = x.P;
    const double left = iz == 0 ? P[p + 1] : P[p - 1];
    const double right = iz == a.nz - 1 ? P[p - 1] : P[p + 1];
    const double D = a.rhs ? a.rhs[p] : 0.0;
    const double c = a.c[ir];
    double residual, diag;
    if (ir == 0) {
      diag = a.e - 2 * c;
      residual = a.a * (right + left) + 4 * c * P[p + a.ldz] + diag * P[p] - D;
    } else {
      diag = a.e;
      residual = a.a * (right + left) + c * P[p + a.ldz] + a.d[ir] * P[p - a.ldz] + a.e * P[p] - D;
    }
    P[p] -= a.omega[x.n] * residual / diag;
    r = fabs(residual);
  }
  x.dump[i] = r;
}

__global__ void sor_state_clear_kernel(int64_t *state) {
  state[0] = 0;
  state[1] = -1;
  state[2] = 0;
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

// src/poisson_correction.c:580-587: Psi[iz][ir] = row[ir]
__global__ __launch_bounds__(256) void line_guess_kernel(double *Psi, int nz, int nr, int ldz, const double *row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nz * nr) return;
  const int ir = (int)(i / nz), iz = (int)(i % nz);
  Psi[fix(ldz, iz, ir)] = row[ir];
}

// src/poisson_correction.c:588-598: the potential of charge *q at z = *centroid,
// r = 0, at (iz*dz + zmin, (ir + 0.5)*dr); sqrt and the division are IEEE
// correctly rounded, like the host's
__global__ __launch_bounds__(256) void point_guess_kernel(double *Psi, int nz, int nr, int ldz, double dz, double dr,
                                                          double zmin, double eps0, const double *q,
                                                          const double *centroid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nz * nr) return;
  const int ir = (int)(i / nz), iz = (int)(i % nz);
  const double a = ((double)iz * dz + zmin) - *centroid, b = dr * ((double)ir + 0.5);
  const double dist = __builtin_sqrt(a * a + b * b);
  Psi[fix(ldz, iz, ir)] = *q / (4 * 3.141592653589793238462643 * eps0 * dist);
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
  state[2] = 0;
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

namespace {

SorCommon sor_common(Run &R, const double *rhs, int oz, int orr) {
  SorCommon a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_dir = R.d.left_bc == SPF_DIRICHLET;
  a.right_dir = R.d.right_bc == SPF_DIRICHLET;
  const double dz = R.d.dz, dr = R.d.dr;
  a.a = 1 / (dz * dz);
  a.e = -2 * (1 / (dz * dz) + 1 / (dr * dr));
  // reciprocals for div_fast, computed with the device's operands
  const double diag0 = a.e - 2 * R.sor_c_host0;
  a.y_e = 1 / a.e;
  a.y_axis = 1 / diag0;
  auto safe = [](double v) {
    double m = std::fabs(v);
    return m > 1e-200 && m < 1e200;
  };
  a.use_y = safe(a.e) && safe(diag0) && safe(a.y_e) && safe(a.y_axis) && a.e < 0 && diag0 < 0;
  a.c = R.sor_c;
  a.d = R.sor_d;
  a.omega = R.omega;
  a.rhs = rhs;
  a.flags = R.flags;
  a.masks = R.sor_masks;
  a.tiles_z = (R.d.nz + 1 + oz - 1) / oz;
  a.ntiles = a.tiles_z * ((R.d.nr + 1 + orr - 1) / orr);
  return a;
}

}  // namespace

// relative margin between summation orders of the stopping rule's sums, each
// of at most nz*(nr-1) non-negative terms
double sor_order_margin(const Run &R) { return order_margin((int64_t)R.d.nz * (R.d.nr - 1), 1.0) - 0x1p-1074; }

namespace {
// calls f with the run's tile shape (an empty SorCfg value)
template <class F>
auto with_shape(const Run &R, F &&f) {
  return R.sor_shape ? f(SorLarge{}) : f(SorSmall{});
}
}  // namespace

int sor_shape_for(int64_t points) { return points >= SPF_SOR_LARGE_POINTS ? 1 : 0; }

int sor_max_batch() { return SorSmall::SK > SorLarge::SK ? SorSmall::SK : SorLarge::SK; }

int64_t sor_tiles(const Run &R) {
  return with_shape(R, [&](auto g) {
    using G = decltype(g);
    int64_t tz = (R.d.nz + 1 + G::OZ - 1) / G::OZ, tr = (R.d.nr + 1 + G::OR - 1) / G::OR;
    return tz * tr;
  });
}

int sor_batch(const Run &R) {
  return with_shape(R, [](auto g) { return decltype(g)::SK; });
}

int64_t sor_mask_words(const Run &R) {
  return with_shape(R, [&](auto g) { return sor_tiles(R) * decltype(g)::NT; });
}

hipError_t launch_sor_masks(Run &R, hipStream_t s) {
  return with_shape(R, [&](auto g) {
    using G = decltype(g);
    SorCommon a = sor_common(R, nullptr, G::OZ, G::OR);
    hipLaunchKernelGGL(sor_masks_kernel<G>, dim3((unsigned)a.ntiles), dim3(G::NT), 0, s, a, R.sor_masks);
    return hipGetLastError();
  });
}

int sor_check_group(const Run &R) {
  return with_shape(R, [](auto g) { return check_group_of<decltype(g)>(); });
}

hipError_t launch_sor_batch(Run &R, const double *Pin, double *Pout, const double *rhs, int64_t n0, int k,
                            int partial_row, int force, hipStream_t s, int64_t check_row, int64_t check_n0,
                            int check_k, double accuracy) {
  return with_shape(R, [&](auto g) {
  using G = decltype(g);
  SorArgs a;
  a.m = sor_common(R, rhs, G::OZ, G::OR);
  a.Pin = Pin;
  a.Pout = Pout;
  a.partials = R.red + (int64_t)partial_row * a.m.ntiles;
  a.state = R.sor_state;
  a.n0 = n0;
  a.k = k;
  a.force = force;
  a.check_partials = R.red + check_row * a.m.ntiles;
  a.sums = R.sor_sums;
  a.accuracy = accuracy;
  a.rel = sor_order_margin(R);
  a.check_n0 = check_n0;
  a.check_k = check_k;
  const unsigned grid = (unsigned)a.m.ntiles + (check_k > 0 ? 1u : 0u);
  hipLaunchKernelGGL(sor_batch_kernel<G>, dim3(grid), dim3(G::NT), 0, s, a);
  return hipGetLastError();
  });
}

// ragged launches over several runs: the per-run table (fixed for a solve),
// then one launch per batch index b
size_t sor_multi_table_bytes(int nruns) {
  return sizeof(SorMultiRun) * (size_t)nruns + sizeof(int32_t) * (size_t)(nruns + 1) + 64;
}

int64_t sor_multi_table(Run *const *runs, const double *const *rhs, double *const (*rings)[6], int nruns,
                        double accuracy, void *host, void *dev, const int32_t **first_dev) {
  SorMultiRun *T = (SorMultiRun *)host;
  int32_t *first = (int32_t *)((char *)host + sizeof(SorMultiRun) * (size_t)nruns);
  int64_t total = 0;
  for (int i = 0; i < nruns; i++) {
    Run &R = *runs[i];
    SorArgs &a = T[i].a;
    memset(&a, 0, sizeof(a));
    a.m = with_shape(R, [&](auto g) { return sor_common(R, rhs[i], decltype(g)::OZ, decltype(g)::OR); });
    a.partials = R.red;
    a.state = R.sor_state;
    a.sums = R.sor_sums;
    a.accuracy = accuracy;
    a.rel = sor_order_margin(R);
    for (int j = 0; j < 6; j++) T[i].ring[j] = rings[i][j];
    first[i] = (int32_t)total;
    total += ((int64_t)a.m.ntiles + 1 + 7) / 8 * 8;
  }
  first[nruns] = (int32_t)total;
  *first_dev = (const int32_t *)((char *)dev + sizeof(SorMultiRun) * (size_t)nruns);
  return total;
}

hipError_t launch_sor_multi(int shape, const void *dev_table, const int32_t *first_dev, int nruns,
                            int64_t total_blocks, int64_t n0, int64_t b, int k, int64_t partial_row, int64_t check_row,
                            int64_t check_n0, int check_k, hipStream_t s) {
  SorMultiLaunch L;
  L.n0 = n0;
  L.b = b;
  L.partial_row = partial_row;
  L.check_row = check_row;
  L.check_n0 = check_n0;
  L.k = k;
  L.check_k = check_k;
  L.nruns = nruns;
  if (shape)
    hipLaunchKernelGGL(sor_multi_kernel<SorLarge>, dim3((unsigned)total_blocks), dim3(SorLarge::NT), 0, s,
                       (const SorMultiRun *)dev_table, first_dev, L);
  else
    hipLaunchKernelGGL(sor_multi_kernel<SorSmall>, dim3((unsigned)total_blocks), dim3(SorSmall::NT), 0, s,
                       (const SorMultiRun *)dev_table, first_dev, L);
  return hipGetLastError();
}

hipError_t launch_sor_check(Run &R, int64_t check_row, int64_t n0, int k, double accuracy, hipStream_t s) {
  CheckArgs c;
  c.partials = R.red + check_row * sor_tiles(R);
  c.ntiles = (int)sor_tiles(R);
  c.k = k;
  c.n0 = n0;
  c.accuracy = accuracy;
  c.rel = sor_order_margin(R);
  c.state = R.sor_state;
  c.sums = R.sor_sums;
  hipLaunchKernelGGL(sor_check_kernel, dim3(1), dim3(CB), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_sor_exact_sweep(Run &R, double *P, const double *rhs, int64_t n, double *dump, hipStream_t s) {
  ExactSweepArgs x;
  x.m = with_shape(R, [&](auto g) { return sor_common(R, rhs, decltype(g)::OZ, decltype(g)::OR); });
  x.P = P;
  x.dump = dump;
  x.n = n;
  const int64_t pts = (int64_t)R.d.nz * (R.d.nr - 1);
  hipLaunchKernelGGL(sor_exact_sweep_kernel, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s, x);
  return hipGetLastError();
}

hipError_t launch_sor_state_clear(Run &R, hipStream_t s) {
  hipLaunchKernelGGL(sor_state_clear_kernel, dim3(1), dim3(1), 0, s, R.sor_state);
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

hipError_t launch_line_guess(Run &R, const double *row, hipStream_t s) {
  const int64_t n = (int64_t)R.d.nz * R.d.nr;
  hipLaunchKernelGGL(line_guess_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.Psi, R.d.nz, R.d.nr,
                     R.ldz, row);
  return hipGetLastError();
}

hipError_t launch_point_guess(Run &R, const double *q, const double *centroid, hipStream_t s) {
  const int64_t n = (int64_t)R.d.nz * R.d.nr;
  hipLaunchKernelGGL(point_guess_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.Psi, R.d.nz, R.d.nr,
                     R.ldz, R.d.dz, R.d.dr, R.d.zmin, R.d.epsilon_o, q, centroid);
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
