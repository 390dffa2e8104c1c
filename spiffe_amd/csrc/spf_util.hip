// spf_util.hip -- deterministic building blocks: fixed-order FP64 reductions.
//
// Every reduction here sums in an order fixed by the problem size alone, so a
// run gives the same bits whatever else shares the GPU.
#include "spf_internal.h"

namespace spf {

namespace {

constexpr int RB = 256;   // reduction block
constexpr int RG = 512;   // reduction grid (fixed)

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

}  // namespace

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
