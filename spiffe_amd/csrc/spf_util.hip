// spf_util.hip -- deterministic building blocks: fixed-order FP64 reductions.
//
// Every reduction here sums in an order fixed by the problem size alone, so a
// run gives the same bits whatever else shares the GPU.
#include "spf_internal.h"

namespace spf {

namespace {

constexpr int RB = 256;   // reduction block
constexpr int RG = 512;   // reduction grid (fixed)

// n terms added one rounded addition at a time from +0 in index order, the
// way a serial loop adds them, by one wavefront: the lanes fetch 64
// consecutive terms while lane 0 adds the previous 64 in order.  The sum is
// valid in lane 0.
template <class T>
__device__ double serial_sum_wave(const T &term, int64_t n, double (*buf)[64]) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  double next = lane < n ? term(lane) : 0.0;
  for (int64_t c = 0; c < n; c += 64) {
    const int par = (int)((c >> 6) & 1);
    buf[par][lane] = next;
    __syncthreads();
    const int64_t i = c + 64 + lane;
    next = i < n ? term(i) : 0.0;
    if (lane == 0) {
      const int m = (int)(n - c < 64 ? n - c : 64);
      for (int k = 0; k < m; k++) s += buf[par][k];
    }
  }
  return s;
}

__device__ double block_sum_fixed(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + off];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

template <bool ABS>
__global__ __launch_bounds__(RB) void region_sum_kernel(const double *a, int ldz, int iz_n, int ir_n,
                                                        double *partials) {
  __shared__ double sh[RB];
  int64_t n = (int64_t)iz_n * ir_n;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)RG * RB) {
    int ir = (int)(i / iz_n), iz = (int)(i % iz_n);
    double v = a[(int64_t)ir * ldz + iz];
    s += ABS ? fabs(v) : v;
  }
  double t = block_sum_fixed(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(RB) void final_sum_kernel(const double *partials, int n, double *out) {
  __shared__ double sh[RB];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += RB) s += partials[i];
  double t = block_sum_fixed(s, sh);
  if (threadIdx.x == 0) *out = t;
}

__global__ void set_i64_kernel(int64_t *p, int64_t v) { *p = v; }

// plain and absolute sums of one region in one pass (fixed order)
__global__ __launch_bounds__(RB) void region_sum2_kernel(const double *a, int ldz, int iz_n, int ir_n,
                                                         double *partials) {
  __shared__ double sh[RB];
  int64_t n = (int64_t)iz_n * ir_n;
  double s = 0, sa = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)RG * RB) {
    int ir = (int)(i / iz_n), iz = (int)(i % iz_n);
    double v = a[(int64_t)ir * ldz + iz];
    s += v;
    sa += fabs(v);
  }
  double t = block_sum_fixed(s, sh);
  double ta = block_sum_fixed(sa, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t;
    partials[RG + blockIdx.x] = ta;
  }
}

// plain and absolute sums of a vector (fixed order)
__global__ __launch_bounds__(RB) void vec_sum2_kernel(const double *a, int64_t n, double *partials) {
  __shared__ double sh[RB];
  double s = 0, sa = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)RG * RB) {
    s += a[i];
    sa += fabs(a[i]);
  }
  double t = block_sum_fixed(s, sh);
  double ta = block_sum_fixed(sa, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t;
    partials[RG + blockIdx.x] = ta;
  }
}

__global__ __launch_bounds__(RB) void final_sum2_kernel(const double *partials, double *out) {
  __shared__ double sh[RB];
  double s = 0, sa = 0;
  for (int i = threadIdx.x; i < RG; i += RB) {
    s += partials[i];
    sa += partials[RG + i];
  }
  double t = block_sum_fixed(s, sh);
  double ta = block_sum_fixed(sa, sh);
  if (threadIdx.x == 0) {
    out[0] = t;
    out[1] = ta;
  }
}

// Terms of a region in the reference's loop orders
struct TermsIzIr {   // for (iz) for (ir): element (iz, ir), optionally |.|
  const double *a;
  int ldz, ir_n, absval;
  __device__ double operator()(int64_t i) const {
    const int iz = (int)(i / ir_n), ir = (int)(i % ir_n);
    const double v = a[(int64_t)ir * ldz + iz];
    return absval ? fabs(v) : v;
  }
};
struct TermsIrIz {   // for (ir) for (iz): element (iz, ir)
  const double *a;
  int ldz, iz_n;
  __device__ double operator()(int64_t i) const {
    const int ir = (int)(i / iz_n), iz = (int)(i % iz_n);
    return a[(int64_t)ir * ldz + iz];
  }
};

struct TermsMomentIrIz {   // for (ir) for (iz): element (iz, ir) times iz*dz + zmin
  const double *a;
  int ldz, iz_n;
  double dz, zmin;
  __device__ double operator()(int64_t i) const {
    const int ir = (int)(i / iz_n), iz = (int)(i % iz_n);
    return a[(int64_t)ir * ldz + iz] * ((double)iz * dz + zmin);
  }
};

template <class T>
__global__ __launch_bounds__(64) void serial_sum_kernel(T term, int64_t n, double *out) {
  __shared__ double buf[2][64];
  const double s = serial_sum_wave(term, n, buf);
  if (threadIdx.x == 0) *out = s;
}

}  // namespace

hipError_t launch_sum2_region(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn,
                              double *out, hipStream_t s) {
  if (pn < 2 * RG) return hipErrorInvalidValue;
  hipLaunchKernelGGL(region_sum2_kernel, dim3(RG), dim3(RB), 0, s, a, ldz, iz_n, ir_n, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(final_sum2_kernel, dim3(1), dim3(RB), 0, s, partials, out);
  return hipGetLastError();
}

hipError_t launch_sum2_vec(const double *a, int64_t n, double *partials, int64_t pn, double *out, hipStream_t s) {
  if (pn < 2 * RG) return hipErrorInvalidValue;
  hipLaunchKernelGGL(vec_sum2_kernel, dim3(RG), dim3(RB), 0, s, a, n, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(final_sum2_kernel, dim3(1), dim3(RB), 0, s, partials, out);
  return hipGetLastError();
}

hipError_t launch_serial_sum_iz_ir(const double *a, int ldz, int iz_n, int ir_n, int absval, double *out,
                                   hipStream_t s) {
  TermsIzIr t{a, ldz, ir_n, absval};
  hipLaunchKernelGGL(serial_sum_kernel<TermsIzIr>, dim3(1), dim3(64), 0, s, t, (int64_t)iz_n * ir_n, out);
  return hipGetLastError();
}

hipError_t launch_serial_moment_ir_iz(const double *a, int ldz, int iz_n, int ir_n, double dz, double zmin, double *out,
                                      hipStream_t s) {
  TermsMomentIrIz t{a, ldz, iz_n, dz, zmin};
  hipLaunchKernelGGL(serial_sum_kernel<TermsMomentIrIz>, dim3(1), dim3(64), 0, s, t, (int64_t)iz_n * ir_n, out);
  return hipGetLastError();
}

hipError_t launch_serial_sum_ir_iz(const double *a, int ldz, int iz_n, int ir_n, double *out, hipStream_t s) {
  TermsIrIz t{a, ldz, iz_n};
  hipLaunchKernelGGL(serial_sum_kernel<TermsIrIz>, dim3(1), dim3(64), 0, s, t, (int64_t)iz_n * ir_n, out);
  return hipGetLastError();
}

// Whether a sum of n terms computed in one order can round to a different
// value than the same terms summed in another order, near a threshold: any
// order of recursive summation of n terms is within g = (n-1)u/(1-(n-1)u)
// times the sum of the terms' magnitudes of the exact sum (u = 2^-53), so two
// orders differ by at most 2g times it.  The returned margin is that bound,
// taken from a computed sum of magnitudes `mag` (itself within g of its exact
// value), with slack for the rounding of this computation.
double order_margin(int64_t n, double mag) {
  const double u = 1.1102230246251565e-16;
  const double k = (double)(n > 1 ? n - 1 : 1);
  const double g = k * u / (1 - k * u);
  return 2.0 * g * mag / (1 - g) * (1 + 1e-12) + 0x1p-1074;
}

hipError_t launch_set_i64(int64_t *p, int64_t v, hipStream_t s) {
  hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_sum_abs(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn, double *out,
                          hipStream_t s) {
  if (pn < RG) return hipErrorInvalidValue;
  hipLaunchKernelGGL(region_sum_kernel<true>, dim3(RG), dim3(RB), 0, s, a, ldz, iz_n, ir_n, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(RB), 0, s, partials, RG, out);
  return hipGetLastError();
}

hipError_t launch_sum_plain(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn, double *out,
                            hipStream_t s) {
  if (pn < RG) return hipErrorInvalidValue;
  hipLaunchKernelGGL(region_sum_kernel<false>, dim3(RG), dim3(RB), 0, s, a, ldz, iz_n, ir_n, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(RB), 0, s, partials, RG, out);
  return hipGetLastError();
}

}  // namespace spf
