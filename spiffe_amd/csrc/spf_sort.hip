// spf_sort.hip -- stable least-significant-digit radix sort of (32-bit key,
// double value) pairs: the ordering step of the deposition engine.
//
// Each pass orders the pairs by one digit of RB bits (RB = 7 or 8; as many
// passes as the key width needs) and keeps the input order among equal
// digits, so after the last pass the pairs are ordered by key and, within a
// key, by input position.  A pass is three launches with no waiting between
// workgroups:
//  1. count: each tile of TILE consecutive pairs counts its digits into
//     cnt[digit][tile];
//  2. scan: one workgroup per digit turns its row into exclusive tile offsets
//     pre[digit][tile] and the digit's total;
//  3. scatter: each tile ranks its pairs by digit in input order (per
//     wavefront, a lane's rank among the lanes that hold the same digit comes
//     from one ballot per digit bit), reorders them by digit in LDS and writes
//     each digit's run to (digit start + tile offset), so the stores of a run
//     are contiguous.
// The count and offset tables are the only state and every pass overwrites
// them; nothing is cleared between calls.
//
// Ranged sort (a device range [kmin, kmax] of the keys other than `discard`):
// every pass orders by digits of rel(k) = k - kmin (discard: kmax - kmin + 1,
// after every real key), 7 bits a pass, and a pass whose digits lie above
// rel's top bit returns at once, so a sort over a narrow range of a wide key
// space costs the passes the range needs.  The passes that ran decide which
// buffer holds the result (sorted_keys); the host never waits for the range.
//
// Device count: the ranged sort can also take the number of pairs from device
// memory (n_dev, at most the host's n).  The tables keep the layout of n's
// tiles, the count and scatter kernels loop over the tiles that hold pairs
// (a grid sized for n, capped), and the scan reads only those tiles.
#include <algorithm>
#include <utility>

#include "spf_internal.h"

namespace spf {

namespace {

constexpr int ST = 256;          // threads per tile workgroup
// pairs per thread, fixed at build time (8, 12 and 16 measured alike)
#ifndef SPF_SORT_ITEMS
#define SPF_SORT_ITEMS 16
#endif
constexpr int SI = SPF_SORT_ITEMS;
constexpr int TILE = ST * SI;    // pairs per tile
constexpr int NW = ST / 64;      // wavefronts per tile workgroup

// the key map of one pass: digits of rel(k) from `shift`; active: the pass runs
struct KeyMap {
  uint32_t kmin, discard_rel, discard;
  bool ranged, active;
  __device__ KeyMap(const uint64_t *range, uint32_t discard_key, int shift) {
    ranged = range != nullptr;
    discard = discard_key;
    kmin = 0;
    discard_rel = 0;
    active = true;
    if (ranged) {
      const uint64_t lo = range[0], hi = range[1];
      const uint32_t span = hi >= lo ? (uint32_t)(hi - lo) + 1u : 0u;   // rel of discard
      kmin = hi >= lo ? (uint32_t)lo : 0u;
      discard_rel = span;
      const int nbits = span ? 32 - __builtin_clz(span) : 0;
      active = shift < nbits;
    }
  }
  __device__ uint32_t rel(uint32_t k) const { return !ranged ? k : (k == discard ? discard_rel : k - kmin); }
};

// pairs present: the host count, or the device count when there is one
__device__ inline int64_t pairs_in(int64_t n, const int64_t *n_dev) {
  if (!n_dev) return n;
  const int64_t m = *n_dev;
  return m < n ? (m > 0 ? m : 0) : n;
}

template <int RB>
__device__ inline void sort_count_tile(const uint32_t *keys, int64_t n, int shift, uint32_t *cnt, int64_t ntiles,
                                       const KeyMap &km, int64_t tile) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t h[R];
  for (int d = threadIdx.x; d < R; d += ST) h[d] = 0;
  __syncthreads();
  const int64_t base = tile * TILE;
  const int lane = threadIdx.x & 63;
  uint32_t k[SI];
#pragma unroll
  for (int i = 0; i < SI; i++) {
    const int64_t idx = base + i * ST + threadIdx.x;
    k[i] = idx < n ? km.rel(keys[idx]) : 0u;
  }
  // After a pass equal digits come in runs, and LDS atomics on one address
  // serialise: the first lane of each run of equal digits adds the run's
  // length.  Valid lanes of a wavefront are a prefix of it.
#pragma unroll
  for (int i = 0; i < SI; i++) {
    const bool valid = base + i * ST + threadIdx.x < n;
    const uint32_t d = (k[i] >> shift) & (R - 1);
    const uint32_t dp = __shfl_up(d, 1, 64);
    const bool start = valid && (lane == 0 || d != dp);
    const unsigned long long starts = __ballot(start);
    const int nv = __popcll(__ballot(valid));
    const unsigned long long above = starts & ~((2ull << lane) - 1ull);
    const int next = above ? __ffsll((long long)above) - 1 : nv;
    if (start) atomicAdd(&h[d], (uint32_t)(next - lane));
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += ST) cnt[(int64_t)d * ntiles + tile] = h[d];
  __syncthreads();   // h is cleared again by the next tile of this workgroup
}

template <int RB>
__global__ __launch_bounds__(ST) void sort_count_kernel(const uint32_t *keys, int64_t n, int shift, uint32_t *cnt,
                                                        int64_t ntiles, const uint64_t *range, uint32_t discard,
                                                        const int64_t *n_dev) {
  const KeyMap km(range, discard, shift);
  if (!km.active) return;
  const int64_t m = pairs_in(n, n_dev), mt = (m + TILE - 1) / TILE;
  for (int64_t t = blockIdx.x; t < mt; t += gridDim.x) sort_count_tile<RB>(keys, m, shift, cnt, ntiles, km, t);
}

// inclusive scans of a and b over the workgroup's threads; returns the
// exclusive values
__device__ inline void block_excl_scan2(uint32_t a, uint32_t b, uint32_t &ea, uint32_t &eb, uint32_t (*ws)[NW]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t ia = a, ib = b;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t va = __shfl_up(ia, off, 64), vb = __shfl_up(ib, off, 64);
    if (lane >= off) {
      ia += va;
      ib += vb;
    }
  }
  if (lane == 63) {
    ws[0][w] = ia;
    ws[1][w] = ib;
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0;
  for (int k = 0; k < w; k++) {
    pa += ws[0][k];
    pb += ws[1][k];
  }
  ea = pa + ia - a;
  eb = pb + ib - b;
}

// one workgroup per digit d: pre[d][t] = sum of cnt[d][t'] over t' < t and
// tot[d] = the sum of the row; thread k scans a contiguous chunk of tiles
__global__ __launch_bounds__(ST) void sort_scan_kernel(const uint32_t *cnt, uint32_t *pre, uint32_t *tot,
                                                       int64_t ntiles, int shift, const uint64_t *range,
                                                       uint32_t discard, int64_t n, const int64_t *n_dev) {
  if (!KeyMap(range, discard, shift).active) return;
  __shared__ uint32_t ws[2][NW];
  const uint32_t *row = cnt + (int64_t)blockIdx.x * ntiles;
  uint32_t *out = pre + (int64_t)blockIdx.x * ntiles;
  const int64_t mt = n_dev ? (pairs_in(n, n_dev) + TILE - 1) / TILE : ntiles;   // tiles holding pairs
  const int64_t per = (mt + ST - 1) / ST;
  const int64_t t0 = std::min<int64_t>((int64_t)threadIdx.x * per, mt), t1 = std::min<int64_t>(t0 + per, mt);
  uint32_t s = 0;
  for (int64_t t = t0; t < t1; t++) s += row[t];
  uint32_t run, unused;
  block_excl_scan2(s, 0u, run, unused, ws);
  for (int64_t t = t0; t < t1; t++) {
    const uint32_t c = row[t];
    out[t] = run;
    run += c;
  }
  if (threadIdx.x == ST - 1) tot[blockIdx.x] = run;
}

// Wavefront w holds the tile's pairs w*SI*64 + j*64 + lane (j = 0..SI-1):
// ranking item j before item j+1 and, within an item, lower lanes first,
// follows the input order.
template <int RB>
__device__ inline void sort_scatter_tile(const uint32_t *kin, const double *vin, uint32_t *kout, double *vout,
                                         int64_t n, int shift, const uint32_t *pre, const uint32_t *tot,
                                         int64_t ntiles, const KeyMap &km, int64_t tile) {
  constexpr int R = 1 << RB;
  static_assert(R <= ST, "one thread per digit");
  __shared__ uint32_t off[NW][R];   // per-wavefront digit counts, then each wavefront's start in the tile
  __shared__ uint32_t gbase[R];     // output position of the digit's run minus its tile-local start
  __shared__ uint32_t ws[2][NW];
  __shared__ uint32_t skey[TILE];
  __shared__ double sval[TILE];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t base = tile * TILE;
  const int nvalid = (int)std::min<int64_t>(TILE, n - base);
  for (int d = threadIdx.x; d < NW * R; d += ST) (&off[0][0])[d] = 0;
  uint32_t key[SI], rank[SI];
  double val[SI];
#pragma unroll
  for (int j = 0; j < SI; j++) {
    const int p = (w * SI + j) * 64 + lane;
    key[j] = p < nvalid ? kin[base + p] : 0u;
    val[j] = p < nvalid ? vin[base + p] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SI; j++) {
    const int p = (w * SI + j) * 64 + lane;
    const uint32_t d = (km.rel(key[j]) >> shift) & (R - 1);
    unsigned long long peers = __ballot(p < nvalid);
#pragma unroll
    for (int b = 0; b < RB; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t below =
        __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
    // every lane reads the count before the lowest lane of each digit group
    // stores the new one (LDS accesses of a wavefront complete in order)
    const uint32_t old = off[w][d];
    rank[j] = old + below;
    if (p < nvalid && below == 0) off[w][d] = old + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // tile count of each digit, wavefront starts, then the digit starts in the
  // tile (exclusive over digits) and in the output (digit totals before it
  // plus the tiles before this one)
  const int d = threadIdx.x;
  uint32_t tcount = 0, gcount = 0;
  if (d < R) {
    uint32_t run = 0;
    for (int k = 0; k < NW; k++) {
      const uint32_t c = off[k][d];
      off[k][d] = run;
      run += c;
    }
    tcount = run;
    gcount = tot[d];
  }
  uint32_t tstart, gstart;
  block_excl_scan2(tcount, gcount, tstart, gstart, ws);
  if (d < R) {
    gbase[d] = gstart + pre[(int64_t)d * ntiles + tile] - tstart;   // modulo 2^32; n < 2^31
    for (int k = 0; k < NW; k++) off[k][d] += tstart;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SI; j++) {
    const int p = (w * SI + j) * 64 + lane;
    if (p < nvalid) {
      const uint32_t pos = off[w][(km.rel(key[j]) >> shift) & (R - 1)] + rank[j];
      skey[pos] = key[j];
      sval[pos] = val[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SI; i++) {
    const int p = i * ST + threadIdx.x;
    if (p < nvalid) {
      const uint32_t k = skey[p];
      const uint32_t dst = gbase[(km.rel(k) >> shift) & (R - 1)] + (uint32_t)p;
      if (dst < (uint32_t)n) {   // always true for consistent tables; keeps stores inside the output
        kout[dst] = k;
        vout[dst] = sval[p];
      }
    }
  }
  __syncthreads();   // the LDS tables are reused by the next tile of this workgroup
}

template <int RB>
__global__ __launch_bounds__(ST) void sort_scatter_kernel(const uint32_t *kin, const double *vin, uint32_t *kout,
                                                          double *vout, int64_t n, int shift, const uint32_t *pre,
                                                          const uint32_t *tot, int64_t ntiles, const uint64_t *range,
                                                          uint32_t discard, const int64_t *n_dev) {
  const KeyMap km(range, discard, shift);
  if (!km.active) return;
  const int64_t m = pairs_in(n, n_dev), mt = (m + TILE - 1) / TILE;
  for (int64_t t = blockIdx.x; t < mt; t += gridDim.x)
    sort_scatter_tile<RB>(kin, vin, kout, vout, m, shift, pre, tot, ntiles, km, t);
}

// workgroups of a count or scatter launch: one per tile, or with a device
// count a capped grid that loops over the tiles holding pairs
constexpr int64_t kDynGrid = 2048;

template <int RB>
hipError_t sort_pass(const uint32_t *ki, const double *vi, uint32_t *ko, double *vo, int64_t n, int shift,
                     uint32_t *cnt, uint32_t *pre, uint32_t *tot, int64_t nt, const uint64_t *range, uint32_t discard,
                     hipStream_t s, const int64_t *n_dev = nullptr) {
  const unsigned grid = (unsigned)(n_dev ? std::min<int64_t>(nt, kDynGrid) : nt);
  hipLaunchKernelGGL(sort_count_kernel<RB>, dim3(grid), dim3(ST), 0, s, ki, n, shift, cnt, nt, range, discard, n_dev);
  hipLaunchKernelGGL(sort_scan_kernel, dim3(1u << RB), dim3(ST), 0, s, cnt, pre, tot, nt, shift, range, discard, n,
                     n_dev);
  hipLaunchKernelGGL(sort_scatter_kernel<RB>, dim3(grid), dim3(ST), 0, s, ki, vi, ko, vo, n, shift, pre, tot, nt,
                     range, discard, n_dev);
  return hipGetLastError();
}

}  // namespace

size_t radix_sort_tmp_bytes(int64_t n) {
  const int64_t nt = (std::max<int64_t>(n, 1) + TILE - 1) / TILE;
  return (size_t)(2 * 256 * nt + 256) * sizeof(uint32_t);
}

hipError_t radix_sort_pairs(uint32_t *k0, double *v0, uint32_t *k1, double *v1, int64_t n, int key_bits, void *tmp,
                            size_t tmp_bytes, uint32_t **kres, double **vres, hipStream_t s) {
  *kres = k0;
  *vres = v0;
  if (n <= 0) return hipSuccess;
  if (n >= (1ll << 31) || key_bits < 1 || key_bits > 32 || tmp_bytes < radix_sort_tmp_bytes(n))
    return hipErrorInvalidValue;
  const int64_t nt = (n + TILE - 1) / TILE;
  const int passes = (key_bits + 7) / 8;
  const int rb = (key_bits + passes - 1) / passes;   // digit width, at most 8
  uint32_t *cnt = static_cast<uint32_t *>(tmp), *pre = cnt + 256 * nt, *tot = pre + 256 * nt;
  uint32_t *ki = k0, *ko = k1;
  double *vi = v0, *vo = v1;
  for (int p = 0; p < passes; p++) {
    // a digit mask wider than rb overlaps the next digit, which the next
    // (stable) pass orders anyway
    const hipError_t e = rb <= 7 ? sort_pass<7>(ki, vi, ko, vo, n, p * rb, cnt, pre, tot, nt, nullptr, 0u, s)
                                 : sort_pass<8>(ki, vi, ko, vo, n, p * rb, cnt, pre, tot, nt, nullptr, 0u, s);
    if (e != hipSuccess) return e;
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  *kres = ki;
  *vres = vi;
  return hipSuccess;
}

hipError_t radix_sort_pairs_ranged(uint32_t *k0, double *v0, uint32_t *k1, double *v1, int64_t n, int key_bits,
                                   const uint64_t *range, uint32_t discard, void *tmp, size_t tmp_bytes,
                                   hipStream_t s, const int64_t *n_dev) {
  if (n <= 0) return hipSuccess;
  if (n >= (1ll << 31) || key_bits < 1 || key_bits > 32 || tmp_bytes < radix_sort_tmp_bytes(n))
    return hipErrorInvalidValue;
  const int64_t nt = (n + TILE - 1) / TILE;
  uint32_t *cnt = static_cast<uint32_t *>(tmp), *pre = cnt + 256 * nt, *tot = pre + 256 * nt;
  uint32_t *ki = k0, *ko = k1;
  double *vi = v0, *vo = v1;
  for (int p = 0; p < sorted_passes_max(key_bits); p++) {
    const hipError_t e = sort_pass<7>(ki, vi, ko, vo, n, 7 * p, cnt, pre, tot, nt, range, discard, s, n_dev);
    if (e != hipSuccess) return e;
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  return hipSuccess;
}

}  // namespace spf
