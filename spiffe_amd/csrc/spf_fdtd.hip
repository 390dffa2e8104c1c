// spf_fdtd.hip -- one fused FDTD leapfrog step of the TM (Er, Ez, Bphi) and
// TE (Ephi, Br, Bz) fields on the staggered r-z Yee mesh, plus the grid
// utility kernels (layout transposes, fills, line currents).
//
// The reference advances each component with its own serial double loop and
// then patches boundary rows (src/integrate.c:277-541).  Every sub-update is a
// pure stencil of the previous sub-update's output, so the final value of each
// array element is a closed function of the time-level-n arrays.  This kernel
// evaluates that function per point.  E and B are ping-ponged (read level n,
// write level n+1).  Each thread computes its point's final E; the B update
// reads the neighbouring new E from LDS or the next lane.  The reference's
// loop `break` after FL_LAST_ACTIVE_PT is reproduced with per-row cut-offs
// (cut1 for loops starting at iz=1, cut0 for iz=0).  Arithmetic repeats the
// reference's operation order; the library is built with -ffp-contract=off so
// no FMA changes a rounding.
#include "spf_internal.h"

namespace spf {

namespace {

constexpr int16_t EZ_ZERO = SPF_FL_EZ_ZERO, ER_ZERO = SPF_FL_ER_ZERO, BPHI_ZERO = SPF_FL_BPHI_ZERO,
                  EPHI_ZERO = SPF_FL_EPHI_ZERO, BZ_ZERO = SPF_FL_BZ_ZERO, BR_ZERO = SPF_FL_BR_ZERO;

// F evaluates the final value of any point with the reference's loop bodies
// and boundary patches.  With NEWE the B updates read the new E from the
// output arrays (already complete) instead of re-evaluating it.
template <bool NEWE>
struct F {
  const FdtdArgs &a;
  __device__ int64_t ix(int iz, int ir) const { return fix(a.ldz, iz, ir); }
  __device__ bool act1(int iz, int ir, int16_t bit) const {
    return !(a.flags[ix(iz, ir)] & bit) && iz <= a.cut1[ir];
  }
  __device__ bool act0(int iz, int ir, int16_t bit) const {
    return !(a.flags[ix(iz, ir)] & bit) && iz <= a.cut0[ir];
  }

  // ---- TM ----
  // Ez loop body, src/integrate.c:341-347 (1 <= iz <= nz-1, 0 <= ir <= nr-1).
  // Every loop body here loads its operands whether or not the point is
  // active, so the loads of a whole call tree issue together; an inactive
  // point returns its old value unchanged.
  __device__ double ez_upd(int iz, int ir) const {
    int64_t p = ix(iz, ir);
    double v = a.Ez[p];
    double inc = a.dti * (-a.Jz[p] / a.eps0 + (a.cp[ir] * a.Bphi[ix(iz, ir + 1)] - a.cm[ir] * a.Bphi[p]));
    return act1(iz, ir, EZ_ZERO) ? v + inc : v;
  }
  // Ez after its loop and the boundary patches of :348-359
  __device__ double ez_final(int iz, int ir) const {
    if (ir == a.nr - 1) return 0.0;
    if (ir == a.nr) return a.Ez[ix(iz, ir)];
    if (iz == 0) {
      double v = ez_upd(1, ir);
      return a.left_neu ? v : -v;
    }
    if (iz == a.nz) {
      double v = ez_upd(a.nz - 1, ir);
      return a.right_neu ? v : -v;
    }
    return ez_upd(iz, ir);
  }
  // Er loop body, :363-372 (0 <= iz <= nz-1, 1 <= ir <= nr-1)
  __device__ double er_upd(int iz, int ir) const {
    int64_t p = ix(iz, ir);
    double v = a.Er[p];
    double inc = a.dti * (-a.Jr[p] / a.eps0 - a.er_factor * (a.Bphi[ix(iz + 1, ir)] - a.Bphi[p]));
    return act0(iz, ir, ER_ZERO) ? v + inc : v;
  }
  // Er after :373-380
  __device__ double er_final(int iz, int ir) const {
    if (iz == a.nz) return a.Er[ix(iz, ir)];
    if (a.left_neu && iz == 0) return 0.0;
    if (a.right_neu && iz == a.nz - 1) return 0.0;
    if (ir == 0) return -er_upd(iz, 1);
    if (ir == a.nr) return a.Er[ix(iz, ir)];
    return er_upd(iz, ir);
  }
  // Bphi loop body, :385-399 (1 <= iz <= nz-1, 1 <= ir <= nr-1), from the final E
  __device__ void bphi_upd(int iz, int ir, double &b, double &b2) const {
    int64_t p = ix(iz, ir);
    const double b0 = a.Bphi[p], b20 = a.Bphi2[p];
    double dB = a.dti * ((ez_at(iz, ir) - ez_at(iz, ir - 1)) / a.dr - (er_at(iz, ir) - er_at(iz - 1, ir)) / a.dz);
    const bool act = act1(iz, ir, BPHI_ZERO);
    b2 = act ? b0 + dB / 2 : b20;
    b = act ? b0 + dB : b0;
  }
  // Bphi after the row patches of :400-408 (1 <= iz <= nz-1, 0 <= ir <= nr)
  __device__ void bphi_inner(int iz, int ir, double &b, double &b2) const {
    if (ir == 0) {
      bphi_upd(iz, 1, b, b2);
      b = -b;
      b2 = -b2;
    } else if (ir == a.nr - 1) {
      bphi_upd(iz, a.nr - 2, b, b2);
      b = b * a.top_factor;
      b2 = b2 * a.top_factor;
    } else if (ir == a.nr) {
      int64_t p = ix(iz, ir);
      b = a.Bphi[p];
      b2 = a.Bphi2[p];
    } else {
      bphi_upd(iz, ir, b, b2);
    }
  }
  // Bphi after the end patches of :409-432
  __device__ void bphi_final(int iz, int ir, double &b, double &b2) const {
    if (iz == 0) {
      bphi_inner(1, ir, b, b2);
      if (!a.left_neu) { b = -b; b2 = -b2; }
    } else if (iz == a.nz) {
      bphi_inner(a.nz - 1, ir, b, b2);
      if (!a.right_neu) { b = -b; b2 = -b2; }
    } else {
      bphi_inner(iz, ir, b, b2);
    }
  }

  // ---- TE ----
  // Ephi loop body, :451-459 (0 <= iz <= nz-1, 1 <= ir <= nr-1)
  __device__ double ephi_upd(int iz, int ir) const {
    int64_t p = ix(iz, ir);
    double v = a.Ephi[p];
    double inc = a.dti * (-a.Jphi[p] / a.eps0 +
                          a.ephi_factor * ((a.Br[ix(iz + 1, ir)] - a.Br[p]) / a.dz - (a.Bz[ix(iz, ir + 1)] - a.Bz[p]) / a.dr));
    return act0(iz, ir, EPHI_ZERO) ? v + inc : v;
  }
  // Ephi after :460-471
  __device__ double ephi_final(int iz, int ir) const {
    if (ir == 0) return 0.0;
    if (a.right_neu && iz == a.nz - 1) return 0.0;
    if (a.left_neu && iz == 0) return 0.0;
    if (a.upper_neu && ir == a.nr - 1) return 0.0;
    if (iz == a.nz || ir == a.nr) return a.Ephi[ix(iz, ir)];
    return ephi_upd(iz, ir);
  }
  // the new E of a point, as the B updates read it
  __device__ double ez_at(int iz, int ir) const { return NEWE ? a.Ezn[ix(iz, ir)] : ez_final(iz, ir); }
  __device__ double er_at(int iz, int ir) const { return NEWE ? a.Ern[ix(iz, ir)] : er_final(iz, ir); }
  __device__ double ephi_at(int iz, int ir) const { return NEWE ? a.Ephin[ix(iz, ir)] : ephi_final(iz, ir); }

  // Br loop body, :474-485 (1 <= iz <= nz-1, 0 <= ir <= nr-1)
  __device__ void br_upd(int iz, int ir, double &b, double &b2) const {
    int64_t p = ix(iz, ir);
    const double b0 = a.Br[p], b20 = a.Br2[p];
    double dB = a.dti * (ephi_at(iz, ir) - ephi_at(iz - 1, ir)) / a.dz;
    const bool act = act1(iz, ir, BR_ZERO);
    b2 = act ? b0 + dB / 2 : b20;
    b = act ? b0 + dB : b0;
  }
  __device__ void br_inner(int iz, int ir, double &b, double &b2) const {
    if (ir == a.nr) {
      int64_t p = ix(iz, ir);
      b = a.Br[p];
      b2 = a.Br2[p];
    } else {
      br_upd(iz, ir, b, b2);
    }
  }
  // Br after :486-509
  __device__ void br_final(int iz, int ir, double &b, double &b2) const {
    if (iz == a.nz) {
      br_inner(a.nz - 1, ir, b, b2);
      if (!a.right_neu) { b = -b; b2 = -b2; }
    } else if (iz == 0) {
      br_inner(1, ir, b, b2);
      if (!a.left_neu) { b = -b; b2 = -b2; }
    } else {
      br_inner(iz, ir, b, b2);
    }
  }
  // Bz loop body, :512-527 (0 <= iz <= nz-1, 1 <= ir <= nr-1)
  __device__ void bz_upd(int iz, int ir, double &b, double &b2) const {
    int64_t p = ix(iz, ir);
    const double b0 = a.Bz[p], b20 = a.Bz2[p];
    double dB = a.bzf[ir] * ((double)(ir + 1) * ephi_at(iz, ir) - (double)ir * ephi_at(iz, ir - 1));
    const bool act = act0(iz, ir, BZ_ZERO);
    b2 = act ? b0 + dB / 2 : b20;
    b = act ? b0 + dB : b0;
  }
  // Bz after :528-539
  __device__ void bz_final(int iz, int ir, double &b, double &b2) const {
    if ((a.right_neu && iz == a.nz - 1) || (a.left_neu && iz == 0)) {
      b = 0.0;
      b2 = 0.0;
      return;
    }
    if (ir == 0) {
      if (iz == a.nz) {
        int64_t p = ix(iz, 1);
        b = a.Bz[p];
        b2 = a.Bz2[p];
      } else {
        bz_upd(iz, 1, b, b2);
      }
      return;
    }
    if (iz == a.nz || ir == a.nr) {
      int64_t p = ix(iz, ir);
      b = a.Bz[p];
      b2 = a.Bz2[p];
      return;
    }
    bz_upd(iz, ir, b, b2);
  }
};

// XCD-aware block order: consecutive hardware workgroup ids go round-robin to
// the 8 XCDs, so hand each XCD a contiguous band of rows; neighbouring rows of
// a band then share that XCD's L2 for their halo reads.
__device__ inline int xcd_band_block(int bid, int nblocks) {
  const int nx = 8;
  int per = (nblocks + nx - 1) / nx;
  int xcd = bid % nx, k = bid / nx;
  int logical = xcd * per + k;
  return logical;
}

// Tiles, two points per thread.  A block is 9 wavefronts x 64 lanes; lane l
// holds columns zs+2l and zs+2l+1 (zs even: each array is read as one 16-byte
// load per lane) and wavefront w row r0-1+w.  Row r0-1 and column zs are halo,
// so the block owns columns zs+1 .. zs+126 (and column 0 in the first tile)
// of rows r0 .. r0+7.  Every thread computes its points' new Ez, Er, Ephi
// into LDS and writes those it owns: points in the E box z in [1, nz-2],
// r in [1, nr-2] with the plain loop bodies (no boundary patch reaches them),
// the rest through F.  Owned points in the B box z in [2, nz-2],
// r in [2, nr-3] (whose E neighbours are all in the E box) then compute their
// B updates from LDS; fdtd_edge_point does the B updates outside it.
constexpr int QZ = 126, QR = 8, QW = QR + 1;

__device__ inline double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ inline double el(const double2 v, int c) { return c ? v.y : v.x; }

__device__ inline bool e_box(const FdtdArgs &a, int iz, int ir) {
  return iz >= 1 && iz <= a.nz - 2 && ir >= 1 && ir <= a.nr - 2;
}
__device__ inline bool b_box(const FdtdArgs &a, int iz, int ir) {
  return iz >= 2 && iz <= a.nz - 2 && ir >= 2 && ir <= a.nr - 3;
}

// The B updates the interior tiles leave: rows 0, 1, nr-2, nr-1, nr and
// columns 0, 1, nz-1, nz, with the boundary patches applied.  They run after
// the tiles, so they read the complete new E.
struct EdgeList {
  int rows[5], nrows, cols[4], ncols;
};

__device__ void fdtd_edge_point(const FdtdArgs &a, const EdgeList &e, int64_t i) {
  const int64_t nrowp = (int64_t)e.nrows * (a.nz + 1);
  const int64_t ncolp = (int64_t)e.ncols * (a.nr + 1);
  if (i >= nrowp + ncolp) return;
  int iz, ir;
  if (i < nrowp) {
    ir = e.rows[i / (a.nz + 1)];
    iz = (int)(i % (a.nz + 1));
  } else {
    const int64_t k = i - nrowp;
    iz = e.cols[k / (a.nr + 1)];
    ir = (int)(k % (a.nr + 1));
    for (int j = 0; j < e.nrows; j++)
      if (e.rows[j] == ir) return;   // already covered as a row point
  }
  F<true> f{a};
  const int64_t q = fix(a.ldz, iz, ir);
  if (a.tm) {
    double b, b2;
    f.bphi_final(iz, ir, b, b2);
    a.Bphin[q] = b;
    a.Bphi2[q] = b2;
  }
  if (a.te) {
    double b, b2, d, d2;
    f.br_final(iz, ir, b, b2);
    f.bz_final(iz, ir, d, d2);
    a.Brn[q] = b;
    a.Br2[q] = b2;
    a.Bzn[q] = d;
    a.Bz2[q] = d2;
  }
}

__global__ __launch_bounds__(256) void fdtd_edge_kernel(FdtdArgs a, EdgeList e) {
  fdtd_edge_point(a, e, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// one tile of the launch (workgroup `bid` of the run's blocks)
__device__ inline void fdtd_tile(const FdtdArgs &a, int tiles_z, int nblocks, int bid) {
  __shared__ double sez[QW][128], ser[QW][128], sephi[QW][128];
  int logical = xcd_band_block(bid, nblocks);
  if (logical >= nblocks) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int zs = (logical % tiles_z) * QZ, r0 = (logical / tiles_z) * QR;
  const int ir = r0 - 1 + w, iz0 = zs + 2 * lane;
  const int ldz = a.ldz;
  const bool rowe = ir >= 1 && ir <= a.nr - 2;   // the row has E-box points
  // clamped so the 16-byte loads stay inside the arrays; values of points
  // outside the E box are never used
  const int izc = min(iz0, ldz - 2), irc = min(max(ir, 0), a.nr - 1);
  const int64_t p = fix(ldz, izc, irc), ph = p + ldz;
  const int fl2 = *reinterpret_cast<const int *>(a.flags + p);   // both columns' flags
  double2 bphi = make_double2(0, 0), br = bphi, bz = bphi;
  if (a.tm) bphi = ld2(a.Bphi + p);
  if (a.te) {
    br = ld2(a.Br + p);
    bz = ld2(a.Bz + p);
  }
  // z+1 neighbours of the odd column through the next lane (lane 63 loads its own)
  double bphi_zp1 = __shfl_down(bphi.x, 1, 64), br_zp1 = __shfl_down(br.x, 1, 64);
  if (lane == 63 && iz0 + 2 <= a.nz) {
    if (a.tm) bphi_zp1 = a.Bphi[fix(ldz, iz0 + 2, irc)];
    if (a.te) br_zp1 = a.Br[fix(ldz, iz0 + 2, irc)];
  }
  double ez_n[2] = {0, 0}, er_n[2] = {0, 0}, ephi_n[2] = {0, 0};
  if (rowe) {
    const int c0 = a.cut0[ir], c1 = a.cut1[ir];
    if (a.tm) {
      const double2 ez = ld2(a.Ez + p), jz = ld2(a.Jz + p), er = ld2(a.Er + p), jr = ld2(a.Jr + p);
      const double2 bphi_hi = ld2(a.Bphi + ph);
      const double cp = a.cp[ir], cm = a.cm[ir];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const int iz = iz0 + c;
        const int16_t fl = (int16_t)(c ? (fl2 >> 16) : (fl2 & 0xffff));
        const double bphi_c = el(bphi, c), bphi_zp = c ? bphi_zp1 : bphi.y;
        ez_n[c] = el(ez, c);
        if (!(fl & EZ_ZERO) && iz <= c1)
          ez_n[c] += a.dti * (-el(jz, c) / a.eps0 + (cp * el(bphi_hi, c) - cm * bphi_c));
        er_n[c] = el(er, c);
        if (!(fl & ER_ZERO) && iz <= c0)
          er_n[c] += a.dti * (-el(jr, c) / a.eps0 - a.er_factor * (bphi_zp - bphi_c));
      }
    }
    if (a.te) {
      const double2 ephi = ld2(a.Ephi + p), jphi = ld2(a.Jphi + p), bz_hi = ld2(a.Bz + ph);
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const int iz = iz0 + c;
        const int16_t fl = (int16_t)(c ? (fl2 >> 16) : (fl2 & 0xffff));
        const double br_c = el(br, c), br_zp = c ? br_zp1 : br.y;
        ephi_n[c] = el(ephi, c);
        if (!(fl & EPHI_ZERO) && iz <= c0)
          ephi_n[c] += a.dti * (-el(jphi, c) / a.eps0 +
                                a.ephi_factor * ((br_zp - br_c) / a.dz - (el(bz_hi, c) - el(bz, c)) / a.dr));
      }
    }
  }
  // points outside the E box (array edges): the loop bodies with the boundary patches
  F<false> f{a};
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int iz = iz0 + c;
    if (ir < 0 || ir > a.nr || iz > a.nz || e_box(a, iz, ir)) continue;
    if (a.tm) {
      ez_n[c] = f.ez_final(iz, ir);
      er_n[c] = f.er_final(iz, ir);
    }
    if (a.te) ephi_n[c] = f.ephi_final(iz, ir);
  }
  reinterpret_cast<double2 *>(sez[w])[lane] = make_double2(ez_n[0], ez_n[1]);
  reinterpret_cast<double2 *>(ser[w])[lane] = make_double2(er_n[0], er_n[1]);
  reinterpret_cast<double2 *>(sephi[w])[lane] = make_double2(ephi_n[0], ephi_n[1]);
  __syncthreads();
  if (w == 0 || ir > a.nr) return;
  // new B of the owned points in the B box; then every owned value is stored,
  // a lane's two columns as one 16-byte store when both are written
  bool own[2], bb[2];
  double bphi_n[2], bphi2_n[2], br_n[2], br2_n[2], bz_n[2], bz2_n[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int iz = iz0 + c, k = 2 * lane + c;
    // owned: columns zs+1 .. zs+126, and column 0 in the first tile
    own[c] = !(iz > a.nz || k > QZ || (k == 0 && zs != 0));
    bb[c] = own[c] && b_box(a, iz, ir);   // implies k >= 1 and an E-box row: the LDS neighbours exist
    bphi_n[c] = bphi2_n[c] = br_n[c] = br2_n[c] = bz_n[c] = bz2_n[c] = 0;
    if (!bb[c]) continue;
    const int64_t q = fix(ldz, iz, ir);
    const int16_t fl = (int16_t)(c ? (fl2 >> 16) : (fl2 & 0xffff));
    if (a.tm) {
      double b = el(bphi, c), b2;
      if (!(fl & BPHI_ZERO) && iz <= a.cut1[ir]) {
        double dB = a.dti * ((ez_n[c] - sez[w - 1][k]) / a.dr - (er_n[c] - ser[w][k - 1]) / a.dz);
        b2 = b + dB / 2;
        b = b + dB;
      } else {
        b2 = a.Bphi2[q];
      }
      bphi_n[c] = b;
      bphi2_n[c] = b2;
    }
    if (a.te) {
      double b = el(br, c), b2;
      if (!(fl & BR_ZERO) && iz <= a.cut1[ir]) {
        double dB = a.dti * (ephi_n[c] - sephi[w][k - 1]) / a.dz;
        b2 = b + dB / 2;
        b = b + dB;
      } else {
        b2 = a.Br2[q];
      }
      double d = el(bz, c), d2;
      if (!(fl & BZ_ZERO) && iz <= a.cut0[ir]) {
        double dB = a.bzf[ir] * ((double)(ir + 1) * ephi_n[c] - (double)ir * sephi[w - 1][k]);
        d2 = d + dB / 2;
        d = d + dB;
      } else {
        d2 = a.Bz2[q];
      }
      br_n[c] = b;
      br2_n[c] = b2;
      bz_n[c] = d;
      bz2_n[c] = d2;
    }
  }
  const int64_t q0 = fix(ldz, min(iz0, ldz - 2), ir);
  auto put = [&](double *dst, const double (&v)[2], const bool (&m)[2]) {
    if (m[0] && m[1])
      *reinterpret_cast<double2 *>(dst + q0) = make_double2(v[0], v[1]);
    else {
      if (m[0]) dst[q0] = v[0];
      if (m[1]) dst[q0 + 1] = v[1];
    }
  };
  if (a.tm) {
    put(a.Ezn, ez_n, own);
    put(a.Ern, er_n, own);
    put(a.Bphin, bphi_n, bb);
    put(a.Bphi2, bphi2_n, bb);
  }
  if (a.te) {
    put(a.Ephin, ephi_n, own);
    put(a.Brn, br_n, bb);
    put(a.Br2, br2_n, bb);
    put(a.Bzn, bz_n, bb);
    put(a.Bz2, bz2_n, bb);
  }
}

__global__ __launch_bounds__(64 * QW) void fdtd_kernel(FdtdArgs a, int tiles_z, int nblocks) {
  fdtd_tile(a, tiles_z, nblocks, (int)blockIdx.x);
}

// Ragged launches over several runs' FDTD steps: workgroups [first[r],
// first[r+1]) are run r's (tile spans rounded to multiples of 8 so the XCD
// banding of each run is the single-run one; edge spans plain).
struct FdtdMultiRun {
  FdtdArgs a;
  EdgeList el;
  int tiles_z, nblocks;
  int64_t ne;   // edge points
};

__device__ inline int ragged_run_f(const int32_t *first, int n, int b) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__global__ __launch_bounds__(64 * QW) void fdtd_multi_kernel(const FdtdMultiRun *__restrict__ T,
                                                              const int32_t *__restrict__ first, int nruns) {
  const int r = ragged_run_f(first, nruns, (int)blockIdx.x);
  const FdtdMultiRun &x = T[r];
  fdtd_tile(x.a, x.tiles_z, x.nblocks, (int)blockIdx.x - first[r]);
}

__global__ __launch_bounds__(256) void fdtd_edge_multi_kernel(const FdtdMultiRun *__restrict__ T,
                                                               const int32_t *__restrict__ first, int nruns) {
  const int r = ragged_run_f(first, nruns, (int)blockIdx.x);
  const FdtdMultiRun &x = T[r];
  fdtd_edge_point(x.a, x.el, (int64_t)((int)blockIdx.x - first[r]) * blockDim.x + threadIdx.x);
}

__global__ void transpose_in_kernel(const double *h, double *d, int nz, int nr, int ldz) {
  // h: [iz][ir] with row length nr+1; d: [ir][iz] with row length ldz
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  if (i >= n) return;
  int ir = (int)(i / (nz + 1)), iz = (int)(i % (nz + 1));
  d[(int64_t)ir * ldz + iz] = h[(int64_t)iz * (nr + 1) + ir];
}

__global__ void transpose_out_kernel(const double *d, double *h, int nz, int nr, int ldz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  if (i >= n) return;
  int iz = (int)(i / (nr + 1)), ir = (int)(i % (nr + 1));
  h[i] = d[(int64_t)ir * ldz + iz];
}

__global__ void transpose_in_i16_kernel(const int16_t *h, int16_t *d, int nz, int nr, int ldz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  if (i >= n) return;
  int ir = (int)(i / (nz + 1)), iz = (int)(i % (nz + 1));
  d[(int64_t)ir * ldz + iz] = h[(int64_t)iz * (nr + 1) + ir];
}

// up to three arrays of n values set to v in one launch (null pointers skipped)
__global__ void fill3_kernel(double *p0, double *p1, double *p2, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * blockDim.x) {
    double *p = i < n ? p0 : (i < 2 * n ? p1 : p2);
    if (p) p[i % n] = v;
  }
}

__global__ void scale_kernel(double *p, int64_t n, double f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = p[i] * f;
}

// src/antenna.c:137-154: the line loop adds the same J to each point in order;
// points are distinct, so one thread per point is the same arithmetic.
__global__ void add_line_kernel(double *J, int ldz, int direction, int64_t ipos, int64_t istart, int64_t iend,
                                double v) {
  int64_t i = istart + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > iend) return;
  int64_t p;
  if (direction == 1)
    p = fix(ldz, (int)ipos, (int)i);  // Jr[ipos][i]
  else
    p = fix(ldz, (int)i, (int)ipos);  // Jz[i][ipos], Jphi[i][ipos]
  J[p] = J[p] + v;
}

}  // namespace

EdgeList edge_list(const FdtdArgs &a) {
  // every row and column outside the interior boxes, each listed once
  EdgeList el;
  el.nrows = el.ncols = 0;
  const int rows[5] = {0, 1, a.nr - 2, a.nr - 1, a.nr}, cols[4] = {0, 1, a.nz - 1, a.nz};
  for (int r : rows) {
    bool dup = r < 0;
    for (int j = 0; j < el.nrows; j++) dup = dup || el.rows[j] == r;
    if (!dup) el.rows[el.nrows++] = r;
  }
  for (int c : cols) {
    bool dup = c < 0;
    for (int j = 0; j < el.ncols; j++) dup = dup || el.cols[j] == c;
    if (!dup) el.cols[el.ncols++] = c;
  }
  return el;
}

hipError_t launch_fdtd(const FdtdArgs &a, hipStream_t s) {
  int tiles_z = (a.nz + 1 + QZ - 1) / QZ, tiles_r = (a.nr + 1 + QR - 1) / QR;
  int nblocks = tiles_z * tiles_r;
  EdgeList el = edge_list(a);
  const int64_t ne = (int64_t)el.nrows * (a.nz + 1) + (int64_t)el.ncols * (a.nr + 1);
  hipLaunchKernelGGL(fdtd_kernel, dim3((nblocks + 7) / 8 * 8), dim3(64 * QW), 0, s, a, tiles_z, nblocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the edge B updates read the complete new E
  hipLaunchKernelGGL(fdtd_edge_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, a, el);
  return hipGetLastError();
}

size_t fdtd_multi_table_bytes(int nruns) {
  return sizeof(FdtdMultiRun) * (size_t)nruns + 2 * sizeof(int32_t) * (size_t)(nruns + 1) + 64;
}

// the tables of one ragged FDTD step: per-run arguments, then the first
// workgroup of each run in the tile launch and in the edge launch
void fdtd_multi_table(const FdtdArgs *args, int nruns, void *host, int64_t *tile_blocks, int64_t *edge_blocks) {
  FdtdMultiRun *T = (FdtdMultiRun *)host;
  int32_t *ft = (int32_t *)((char *)host + sizeof(FdtdMultiRun) * (size_t)nruns);
  int32_t *fe = ft + nruns + 1;
  int64_t nt = 0, ne = 0;
  for (int i = 0; i < nruns; i++) {
    FdtdMultiRun &x = T[i];
    x.a = args[i];
    x.tiles_z = (args[i].nz + 1 + QZ - 1) / QZ;
    x.nblocks = x.tiles_z * ((args[i].nr + 1 + QR - 1) / QR);
    x.el = edge_list(args[i]);
    x.ne = (int64_t)x.el.nrows * (args[i].nz + 1) + (int64_t)x.el.ncols * (args[i].nr + 1);
    ft[i] = (int32_t)nt;
    fe[i] = (int32_t)ne;
    nt += ((int64_t)x.nblocks + 7) / 8 * 8;
    ne += (x.ne + 255) / 256;
  }
  ft[nruns] = (int32_t)nt;
  fe[nruns] = (int32_t)ne;
  *tile_blocks = nt;
  *edge_blocks = ne;
}

hipError_t launch_fdtd_multi(const void *dev_table, int nruns, int64_t tile_blocks, int64_t edge_blocks,
                             hipStream_t s) {
  const FdtdMultiRun *T = (const FdtdMultiRun *)dev_table;
  const int32_t *ft = (const int32_t *)((const char *)dev_table + sizeof(FdtdMultiRun) * (size_t)nruns);
  hipLaunchKernelGGL(fdtd_multi_kernel, dim3((unsigned)tile_blocks), dim3(64 * QW), 0, s, T, ft, nruns);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fdtd_edge_multi_kernel, dim3((unsigned)edge_blocks), dim3(256), 0, s, T, ft + nruns + 1, nruns);
  return hipGetLastError();
}

hipError_t launch_transpose_in(const double *h, double *d, int nz, int nr, int ldz, hipStream_t s) {
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  hipLaunchKernelGGL(transpose_in_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h, d, nz, nr, ldz);
  return hipGetLastError();
}

hipError_t launch_transpose_out(const double *d, double *h, int nz, int nr, int ldz, hipStream_t s) {
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  hipLaunchKernelGGL(transpose_out_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, h, nz, nr, ldz);
  return hipGetLastError();
}

hipError_t launch_transpose_in_i16(const int16_t *h, int16_t *d, int nz, int nr, int ldz, hipStream_t s) {
  int64_t n = (int64_t)(nz + 1) * (nr + 1);
  hipLaunchKernelGGL(transpose_in_i16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h, d, nz, nr, ldz);
  return hipGetLastError();
}

hipError_t launch_fill3(double *p0, double *p1, double *p2, int64_t n, double v, hipStream_t s) {
  if (n <= 0 || (!p0 && !p1 && !p2)) return hipSuccess;
  int64_t blocks = (3 * n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill3_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p0, p1, p2, n, v);
  return hipGetLastError();
}

hipError_t launch_scale(double *p, int64_t n, double f, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, f);
  return hipGetLastError();
}

hipError_t launch_add_line(double *J, int ldz, int direction, int64_t ipos, int64_t istart, int64_t iend, double v,
                           hipStream_t s) {
  if (iend < istart) return hipSuccess;
  int64_t n = iend - istart + 1;
  hipLaunchKernelGGL(add_line_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, J, ldz, direction, ipos,
                     istart, iend, v);
  return hipGetLastError();
}

}  // namespace spf
