// spf_util.hip -- deterministic building blocks: ordered segment-start
// compaction (position order preserved) and fixed-order FP64 reductions.
//
// Every reduction here sums in an order fixed by the problem size alone, so a
// run gives the same bits whatever else shares the GPU.
#include "spf_internal.h"

namespace spf {

namespace {

constexpr int CB = 256;       // threads per compaction block
constexpr int CITEMS = 16;    // items per thread
constexpr int CSPAN = CB * CITEMS;

__device__ int block_scan_flag(bool f, int *wsum, int &excl) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long m = __ballot(f);
  unsigned long long lt = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int before = 0, total = 0;
  for (int k = 0; k < nw; k++) {
    int v = wsum[k];
    if (k < w) before += v;
    total += v;
  }
  __syncthreads();
  excl = before + __popcll(lt);
  return total;
}

// Position i of sorted keys starts a segment when its key differs from the
// previous one and is not the discard key.
struct SegStart {
  const uint32_t *sk;
  uint32_t discard;
  __device__ bool operator()(int64_t i) const {
    const uint32_t k = sk[i];
    return k != discard && (i == 0 || sk[i - 1] != k);
  }
};

// Per-block segment counts; also records in counters[6] the first position
// holding the discard key (counters[6] is preset to n).
__global__ __launch_bounds__(CB) void seg_count_kernel(SegStart f, int64_t n, int32_t *bcount, int64_t *counters) {
  __shared__ int wsum[CB / 64];
  const int64_t base = (int64_t)blockIdx.x * CSPAN;
  int c = 0;
  for (int j = 0; j < CITEMS; j++) {
    const int64_t i = base + (int64_t)j * CB + threadIdx.x;
    if (i < n) {
      if (f(i)) c++;
      if (f.sk[i] == f.discard && (i == 0 || f.sk[i - 1] != f.discard)) counters[6] = i;
    }
  }
  // wave sums then block sum (integer: order irrelevant)
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < CB / 64; k++) t += wsum[k];
    bcount[blockIdx.x] = t;
  }
}

// single workgroup: exclusive scan of nb block counts, total into *count
__global__ __launch_bounds__(1024) void block_scan_kernel(const int32_t *bcount, int64_t nb, int32_t *boff,
                                                          int64_t *count) {
  __shared__ int buf[1024];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t c = 0; c < nb; c += 1024) {
    const int64_t k = c + threadIdx.x;
    const int v = k < nb ? bcount[k] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    const int incl = buf[threadIdx.x];
    if (k < nb) boff[k] = (int32_t)(carry + incl - v);
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

// segment starts in position order at their block's offset
__global__ __launch_bounds__(CB) void seg_scatter_kernel(SegStart f, int64_t n, const int32_t *boff, int32_t *out) {
  __shared__ int wsum[CB / 64];
  const int64_t base = (int64_t)blockIdx.x * CSPAN;
  int running = boff[blockIdx.x];
  for (int j = 0; j < CITEMS; j++) {
    const int64_t i = base + (int64_t)j * CB + threadIdx.x;
    const bool fl = i < n && f(i);
    int excl;
    const int tot = block_scan_flag(fl, wsum, excl);
    if (fl) out[running + excl] = (int32_t)i;
    running += tot;
  }
}

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

int64_t compact_blocks(int64_t n) { return (n + CSPAN - 1) / CSPAN; }

hipError_t segment_starts(const uint32_t *sk, int64_t n, uint32_t discard, int32_t *out, int64_t *counters,
                          int32_t *bcount, int32_t *boff, hipStream_t s) {
  const int64_t nb = compact_blocks(n);
  if (nb == 0) return launch_set_i64(counters + 5, 0, s);
  SegStart f{sk, discard};
  hipLaunchKernelGGL(seg_count_kernel, dim3((unsigned)nb), dim3(CB), 0, s, f, n, bcount, counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(block_scan_kernel, dim3(1), dim3(1024), 0, s, bcount, nb, boff, counters + 5);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(seg_scatter_kernel, dim3((unsigned)nb), dim3(CB), 0, s, f, n, boff, out);
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
