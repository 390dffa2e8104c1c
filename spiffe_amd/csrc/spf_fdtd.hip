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
  // Ez loop body, src/integrate.c:341-347 (1 <= iz <= nz-1, 0 <= ir <= nr-1)
  __device__ double ez_upd(int iz, int ir) const {
    int64_t p = ix(iz, ir);
    double v = a.Ez[p];
    if (act1(iz, ir, EZ_ZERO))
      v += a.dti * (-a.Jz[p] / a.eps0 + (a.cp[ir] * a.Bphi[ix(iz, ir + 1)] - a.cm[ir] * a.Bphi[p]));
    return v;
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
    if (act0(iz, ir, ER_ZERO))
      v += a.dti * (-a.Jr[p] / a.eps0 - a.er_factor * (a.Bphi[ix(iz + 1, ir)] - a.Bphi[p]));
    return v;
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
    b = a.Bphi[p];
    b2 = a.Bphi2[p];
    if (act1(iz, ir, BPHI_ZERO)) {
      double dB = a.dti * ((ez_final(iz, ir) - ez_final(iz, ir - 1)) / a.dr - (er_final(iz, ir) - er_final(iz - 1, ir)) / a.dz);
      b2 = b + dB / 2;
      b = b + dB;
    }
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
    if (act0(iz, ir, EPHI_ZERO))
      v += a.dti * (-a.Jphi[p] / a.eps0 +
                    a.ephi_factor * ((a.Br[ix(iz + 1, ir)] - a.Br[p]) / a.dz - (a.Bz[ix(iz, ir + 1)] - a.Bz[p]) / a.dr));
    return v;
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
  // Br loop body, :474-485 (1 <= iz <= nz-1, 0 <= ir <= nr-1)
  __device__ void br_upd(int iz, int ir, double &b, double &b2) const {
    int64_t p = ix(iz, ir);
    b = a.Br[p];
    b2 = a.Br2[p];
    if (act1(iz, ir, BR_ZERO)) {
      double dB = a.dti * (ephi_final(iz, ir) - ephi_final(iz - 1, ir)) / a.dz;
      b2 = b + dB / 2;
      b = b + dB;
    }
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
    b = a.Bz[p];
    b2 = a.Bz2[p];
    if (act0(iz, ir, BZ_ZERO)) {
      double dB = a.bzf[ir] * ((double)(ir + 1) * ephi_final(iz, ir) - (double)ir * ephi_final(iz, ir - 1));
      b2 = b + dB / 2;
      b = b + dB;
    }
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

// One point per thread.  A block is 9 wavefronts x 64 lanes: wavefront w is
// row r0-1+w and lane l column z0-1+l, so row r0-1 and column z0-1 are halo
// and the block outputs a 63 x 8 tile.  Every thread computes its point's
// final Ez, Er, Ephi into LDS; then each output thread computes its B updates.
// Points in the box z in [1, nz-2], r in [1, nr-2] use the plain loop bodies
// (no boundary patch reaches them); all other points go through F, which
// applies the patches.  The B updates of points in z in [2, nz-2],
// r in [2, nr-3] read their neighbours' new E from LDS; the others use F too.
constexpr int QZ = 63, QR = 8, QW = QR + 1;

__global__ __launch_bounds__(64 * QW) void fdtd_kernel(FdtdArgs a, int tiles_z, int nblocks) {
  __shared__ double sez[QW][64], ser[QW][64], sephi[QW][64];
  int logical = xcd_band_block(blockIdx.x, nblocks);
  if (logical >= nblocks) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int z0 = (logical % tiles_z) * QZ, r0 = (logical / tiles_z) * QR;
  const int iz = z0 - 1 + lane, ir = r0 - 1 + w;
  const bool inside = iz >= 0 && iz <= a.nz && ir >= 0 && ir <= a.nr;
  const int ldz = a.ldz;
  const int izc = min(max(iz, 0), a.nz), irc = min(max(ir, 0), a.nr);
  const int64_t p = fix(ldz, izc, irc);
  F f{a};
  const bool efast = iz >= 1 && iz <= a.nz - 2 && ir >= 1 && ir <= a.nr - 2;
  const int16_t fl = a.flags[p];
  // neighbours z+1 through the lane above (lane 63 loads its own)
  double bphi_c = 0, br_c = 0;
  if (a.tm) bphi_c = a.Bphi[p];
  if (a.te) br_c = a.Br[p];
  double bphi_zp = __shfl_down(bphi_c, 1, 64), br_zp = __shfl_down(br_c, 1, 64);
  if (lane == 63 && efast) {
    if (a.tm) bphi_zp = a.Bphi[p + 1];
    if (a.te) br_zp = a.Br[p + 1];
  }
  double ez_n = 0, er_n = 0, ephi_n = 0;
  if (inside) {
    if (efast) {
      const int c0 = a.cut0[ir], c1 = a.cut1[ir];
      if (a.tm) {
        const double bphi_hi = a.Bphi[p + ldz];
        ez_n = a.Ez[p];
        if (!(fl & EZ_ZERO) && iz <= c1)
          ez_n += a.dti * (-a.Jz[p] / a.eps0 + (a.cp[ir] * bphi_hi - a.cm[ir] * bphi_c));
        er_n = a.Er[p];
        if (!(fl & ER_ZERO) && iz <= c0)
          er_n += a.dti * (-a.Jr[p] / a.eps0 - a.er_factor * (bphi_zp - bphi_c));
      }
      if (a.te) {
        const double bz_c = a.Bz[p], bz_hi = a.Bz[p + ldz];
        ephi_n = a.Ephi[p];
        if (!(fl & EPHI_ZERO) && iz <= c0)
          ephi_n += a.dti * (-a.Jphi[p] / a.eps0 + a.ephi_factor * ((br_zp - br_c) / a.dz - (bz_hi - bz_c) / a.dr));
      }
    } else {
      if (a.tm) {
        ez_n = f.ez_final(iz, ir);
        er_n = f.er_final(iz, ir);
      }
      if (a.te) ephi_n = f.ephi_final(iz, ir);
    }
  }
  sez[w][lane] = ez_n;
  ser[w][lane] = er_n;
  sephi[w][lane] = ephi_n;
  __syncthreads();
  if (!inside || w == 0 || lane == 0 || iz >= z0 + QZ) return;
  const bool bfast = iz >= 2 && iz <= a.nz - 2 && ir >= 2 && ir <= a.nr - 3;
  if (a.tm) {
    double b, b2;
    if (bfast) {
      b = bphi_c;
      if (!(fl & BPHI_ZERO) && iz <= a.cut1[ir]) {
        double dB = a.dti * ((ez_n - sez[w - 1][lane]) / a.dr - (er_n - ser[w][lane - 1]) / a.dz);
        b2 = b + dB / 2;
        b = b + dB;
      } else {
        b2 = a.Bphi2[p];
      }
    } else {
      f.bphi_final(iz, ir, b, b2);
    }
    a.Ezn[p] = ez_n;
    a.Ern[p] = er_n;
    a.Bphin[p] = b;
    a.Bphi2[p] = b2;
  }
  if (a.te) {
    double b, b2, c, c2;
    if (bfast) {
      b = br_c;
      if (!(fl & BR_ZERO) && iz <= a.cut1[ir]) {
        double dB = a.dti * (ephi_n - sephi[w][lane - 1]) / a.dz;
        b2 = b + dB / 2;
        b = b + dB;
      } else {
        b2 = a.Br2[p];
      }
      c = a.Bz[p];
      if (!(fl & BZ_ZERO) && iz <= a.cut0[ir]) {
        double dB = a.bzf[ir] * ((double)(ir + 1) * ephi_n - (double)ir * sephi[w - 1][lane]);
        c2 = c + dB / 2;
        c = c + dB;
      } else {
        c2 = a.Bz2[p];
      }
    } else {
      f.br_final(iz, ir, b, b2);
      f.bz_final(iz, ir, c, c2);
    }
    a.Ephin[p] = ephi_n;
    a.Brn[p] = b;
    a.Br2[p] = b2;
    a.Bzn[p] = c;
    a.Bz2[p] = c2;
  }
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

__global__ void fill_kernel(double *p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
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

hipError_t launch_fdtd(const FdtdArgs &a, hipStream_t s) {
  int tiles_z = (a.nz + 1 + QZ - 1) / QZ, tiles_r = (a.nr + 1 + QR - 1) / QR;
  int nblocks = tiles_z * tiles_r;
  int grid = ((nblocks + 7) / 8) * 8;
  hipLaunchKernelGGL(fdtd_kernel, dim3(grid), dim3(64 * QW), 0, s, a, tiles_z, nblocks);
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

hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
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
