// spf_deposit.hip -- nearest-grid-point current deposition
// (src/space_charge.c:34-295) and the particle-charge removal of the Poisson
// correction (src/poisson_correction.c:184-243, :377-437), with the
// reference's per-cell summation order and no floating-point atomics.
//
// The reference adds each particle's contribution to its cell in particle
// index order, one pass per component.  Floating-point addition is not
// associative, so the device keeps exactly that order:
//  1. key: for every active particle and component c the target cell and the
//     contribution value, key = cell * 4 + c, value = the contribution; the
//     kernel also finds the smallest and largest key in use;
//  2. a stable radix sort of the (key, value) pairs over that key range (a
//     bunch covers a band of grid rows, so fewer 7-bit passes cover it than
//     the whole key space would need): inputs
//     are laid out in particle index order per component, so each cell's
//     contributions come out contiguous and still in index order;
//  3. each cell segment (the run of one key) is folded sequentially onto the
//     cell's current value by one thread, or by one wavefront when it is
//     longer than LONG_SEG values.
// Every cell has exactly one writer, so the result is the reference's bits.
// Contributions that are exactly zero in the reference (ir >= nr-1 for Jz and
// Jphi, ir == 0 and ir == nr-1 for Jr) are left out: J starts each step at +0
// and a running sum of nonzero doubles from +0 is never -0, so adding a zero
// cannot change a bit.  The charge removal adds -Q, which equals subtracting Q.
#include <algorithm>
#include <cstring>

#include "spf_internal.h"

namespace spf {

namespace {

constexpr int COMP_JZ = 0, COMP_JR = 1, COMP_JPHI = 2, COMP_Q = 3;

struct KeyArgs {
  int nz, nr, ncomp, ldz;
  int comp[4];                  // component id of slot k
  double zmin, dz, dr, one_over_piDzDr2;
  const double *Q, *z, *r, *vz, *vr, *vphi;
  const uint32_t *status;
  const int16_t *flags;
  int left_neu, right_neu, upper_neu;
  const int64_t *np;
  int64_t n_ub;
  uint32_t *keys;
  double *vals;
  uint32_t discard;
  int cell_bits;
  int64_t *counters;            // [9] (number of long segments) cleared by thread 0
  unsigned long long *krange;   // [0, 256): per-slot smallest key, [256, 512): largest, [512], [513]: the range
  double *clear[3];             // target arrays to set to +0 first (or null), clear_n cells each
  int64_t clear_n;
  int zero_terms;               // COMP_Q: apply the signed-zero effect of the zero-weight terms
  double *zero_target;          // the Q array those terms act on
};

// NGP value of the Jz/Jphi loops for cell row ir (0 <= ir <= nr-2), :92-100, :254-262
__device__ inline double jm_centered(double QVG, long long ir) {
  if (ir == 0) return 4.0 * 1.0 * QVG;
  return 1.0 * QVG / ((ir + 0.5) * (ir + 0.5) - (ir - 0.5) * (ir - 0.5));
}

// The key range: each key workgroup folds its keys' smallest and largest into
// one of 256 slots (integer min/max: the same result in any order; the slots
// keep the atomics off one address), and one workgroup then folds the slots.
constexpr int kRangeSlots = 256;
__global__ __launch_bounds__(kRangeSlots) void key_range_reset_kernel(unsigned long long *krange) {
  krange[threadIdx.x] = ~0ull;
  krange[kRangeSlots + threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kRangeSlots) void key_range_final_kernel(unsigned long long *krange) {
  __shared__ unsigned long long lo[kRangeSlots], hi[kRangeSlots];
  lo[threadIdx.x] = krange[threadIdx.x];
  hi[threadIdx.x] = krange[kRangeSlots + threadIdx.x];
  __syncthreads();
  for (int off = kRangeSlots / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + off]);
      hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    krange[2 * kRangeSlots] = lo[0];
    krange[2 * kRangeSlots + 1] = hi[0];
  }
}

__global__ __launch_bounds__(256) void key_kernel(KeyArgs a) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip == 0) a.counters[9] = 0;
  // the targets' reset (src/integrate.c:169 zeroes J before the deposition),
  // done here to save a launch; the folds that add into them come later
  for (int64_t q = ip; q < a.clear_n; q += (int64_t)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; c++)
      if (a.clear[c]) a.clear[c][q] = 0.0;
  // smallest and largest key this lane writes (every lane takes part in the
  // wavefront's reduction below, so no lane returns early)
  uint32_t kmin = 0xffffffffu, kmax = 0u;
  const bool in = ip < a.n_ub;
  int64_t np = in ? *a.np : 0;
  bool live = in && ip < np && (a.status[ip] & SPF_PART_ACTIVE);
  long long izh = -1, izA = -1, irh = -1, irA = -1;
  double Qo = 0;
  if (live) {
    double x = (a.z[ip] - a.zmin) / a.dz;
    double y = a.r[ip] / a.dr;
    double th = x + 0.5, t1 = x + 1.0, yh = y + 0.5, y1 = y + 1.0;
    // every valid cell index lies in [0, nz] x [0, nr]; outside these
    // bounds (and for NaN) the conversions are skipped
    if (th > -1.0 && th < a.nz + 1.0) izh = (long long)th;
    if (t1 > -1.0 && t1 < a.nz + 1.0) izA = (long long)t1;
    if (yh > -1.0 && yh < a.nr + 1.0) irh = (long long)yh;
    if (y1 > -1.0 && y1 < a.nr + 1.0) irA = (long long)y1;
    Qo = a.Q[ip] * a.one_over_piDzDr2;
  }
  const int nz = a.nz, nr = a.nr;
  for (int k = 0; in && k < a.ncomp; k++) {
    int c = a.comp[k];
    long long iz = -1, ir = -1;
    double v = 0;
    if (live) {
      if (c == COMP_JZ) {
        // iz = x+1.0, ir = y+0.5; valid iz <= nz, ir <= nr; nonzero only for ir <= nr-2
        if (izA >= 0 && izA <= nz && irh >= 0 && irh <= nr - 2) {
          iz = izA;
          ir = irh;
          v = jm_centered(a.vz[ip] * Qo, irh);
        }
      } else if (c == COMP_JR) {
        // iz = x+0.5, ir = y+1.0; valid ir <= nr-1; nonzero only for 1 <= ir <= nr-2
        if (izh >= 0 && izh <= nz && irA >= 1 && irA <= nr - 2) {
          iz = izh;
          ir = irA;
          double QVG = a.vr[ip] * Qo;
          v = 1.0 * QVG / ((double)irA * (double)irA - (double)(irA - 1) * (double)(irA - 1));
        }
      } else if (c == COMP_JPHI) {
        if (izh >= 0 && izh <= nz && irh >= 0 && irh <= nr - 2) {
          iz = izh;
          ir = irh;
          v = jm_centered(a.vphi[ip] * Qo, irh);
        }
      } else {  // COMP_Q: iz = x+0.5, ir = y+0.5 in [0,nz-1] x [0,nr-1], not metal or a Neumann edge
        if (izh >= 0 && izh < nz && irh >= 0 && irh < nr) {
          int16_t f = a.flags[fix(a.ldz, (int)izh, (int)irh)];
          bool blocked = (f & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (izh == 0 && a.left_neu) ||
                         (izh >= nz - 1 && a.right_neu) || (irh >= nr - 1 && a.upper_neu);
          if (!blocked) {
            iz = izh;
            ir = irh;
            v = -(a.Q[ip] * 1.0 * 1.0);
          }
          // The reference also subtracts Q*fb*fr, Q*fa*fl and Q*fa*fr from the
          // cells (iz+1, ir), (iz, ir+1), (iz+1, ir+1) (src/poisson_correction.c:
          // 222-236); without smearing these are Q*0, a zero with the sign of Q.
          // Subtracting +0 changes no value; subtracting -0 (Q < 0) turns a -0
          // cell into +0 and leaves every other value alone.  A cell can hold -0
          // only while nothing but -0 has been added to a -0 start, so turning
          // -0 into +0 before the ordered fold gives the reference's bits
          // whatever the position of the zero terms in its sequence.
          if (a.zero_terms && __builtin_signbit(a.Q[ip])) {
            const long long zc[3][2] = {{izh + 1, irh}, {izh, irh + 1}, {izh + 1, irh + 1}};
            const bool zb[3] = {(izh == 0 && a.left_neu) || (izh >= nz && a.right_neu) || (irh >= nr - 1 && a.upper_neu),
                                (izh == 0 && a.left_neu) || (izh >= nz - 1 && a.right_neu) || (irh >= nr && a.upper_neu),
                                (izh == 0 && a.left_neu) || (izh >= nz && a.right_neu) || (irh >= nr && a.upper_neu)};
#pragma unroll
            for (int t = 0; t < 3; t++) {
              const int64_t pc = fix(a.ldz, (int)zc[t][0], (int)zc[t][1]);
              if (zb[t] || (a.flags[pc] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) continue;
              double *q = a.zero_target + pc;
              if (__double_as_longlong(*q) == (long long)0x8000000000000000ull) *q = 0.0;
            }
          }
        }
      }
    }
    int64_t slot = (int64_t)k * a.n_ub + ip;
    if (iz >= 0) {
      const uint32_t cell = (uint32_t)(ir * (nz + 1) + iz);
      const uint32_t key = (cell << 2) | (uint32_t)k;
      a.keys[slot] = key;
      a.vals[slot] = v;
      kmin = min(kmin, key);
      kmax = max(kmax, key);
    } else {
      a.keys[slot] = a.discard;
      a.vals[slot] = 0.0;
    }
  }
  if (!a.krange) return;   // the sort covers the whole key space
  for (int off = 32; off > 0; off >>= 1) {
    kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, off, 64));
    kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, off, 64));
  }
  __shared__ uint32_t wmin[4], wmax[4];
  if ((threadIdx.x & 63) == 0) {
    wmin[threadIdx.x >> 6] = kmin;
    wmax[threadIdx.x >> 6] = kmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      kmin = min(kmin, wmin[w]);
      kmax = max(kmax, wmax[w]);
    }
    if (kmin <= kmax) {
      const int slot = blockIdx.x % kRangeSlots;
      atomicMin(a.krange + slot, (unsigned long long)kmin);
      atomicMax(a.krange + kRangeSlots + slot, (unsigned long long)kmax);
    }
  }
}

struct FoldArgs {
  int nz, ldz, cell_bits;
  int64_t n;
  uint32_t discard;
  const uint32_t *sk;           // sorted pairs: the sort's first buffer, or sk1/sv1 (ranged_sort_in_k1)
  const double *sv;
  const uint32_t *sk1;
  const double *sv1;
  const uint64_t *range;        // key range of the ranged sort (null: sk/sv hold the result)
  int64_t *counters;    // [9]: number of long segments
  int32_t *long_list;   // first positions of segments longer than LONG_SEG
  double *target[4];
};


__device__ inline double *target_cell(const FoldArgs &a, uint32_t key) {
  uint32_t k = key & 3u;
  uint32_t cell = key >> 2;
  int iz = (int)(cell % (uint32_t)(a.nz + 1)), ir = (int)(cell / (uint32_t)(a.nz + 1));
  return a.target[k] + fix(a.ldz, iz, ir);
}

// The additions of a segment (the run of one key in the sorted pairs) form
// one dependent chain.  Each wavefront looks at 64 consecutive positions: a
// run starts where the key differs from the previous one, and a run that
// starts in the window ends at the next start in it; only the window's last
// run can go on past it, and the wavefront finds where.  The lane of a
// segment's start folds it, from register blocks of 16 values with the next
// block's loads already in flight, unless the key at start + LONG_SEG is still
// the same: such segments are listed and folded by one wavefront each
// (fold_long_kernel).  The discard key's run is not folded.
constexpr int FB = 16;
constexpr int64_t LONG_SEG = 256;

__global__ __launch_bounds__(256) void fold_kernel(FoldArgs a) {
  // the buffer the ranged sort left its result in
  const bool in1 = a.range && ranged_sort_in_k1(a.range);
  const uint32_t *__restrict__ sk = in1 ? a.sk1 : a.sk;
  const double *__restrict__ sv = in1 ? a.sv1 : a.sv;
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, base = p - lane;
  if (base >= a.n) return;
  const bool valid = p < a.n;
  const uint32_t key = valid ? sk[p] : 0u;
  // a run starts here, or the array has ended
  const bool edge = !valid || p == 0 || sk[p - 1] != key;
  const unsigned long long em = __ballot(edge);
  if (em == 0) return;   // the window lies inside one run
  const unsigned long long above = em & ~((2ull << lane) - 1ull);
  int64_t e0 = above ? base + (__ffsll((long long)above) - 1) : 0;
  // the window's last run (from lane L) may go on past the window
  const int L = 63 - __clzll((long long)em);
  const int64_t sL = base + L;
  const uint32_t kL = __shfl(key, L, 64);
  bool longL = false;
  int64_t eL = 0;
  if (sL < a.n && kL != a.discard) {
    if (sL + LONG_SEG < a.n && sk[sL + LONG_SEG] == kL) {
      longL = true;
    } else {
      // sorted keys: the run ends within LONG_SEG positions of sL
      for (int64_t q = base + 64;; q += 64) {
        const int64_t i = q + lane;
        const unsigned long long dm = __ballot(i >= a.n || sk[i] != kL);
        if (dm) {
          eL = q + (__ffsll((long long)dm) - 1);
          break;
        }
      }
    }
  }
  if (!valid || !edge || key == a.discard) return;
  const int64_t s0 = p;
  if (lane == L) {
    if (longL) {
      const unsigned long long slot = atomicAdd((unsigned long long *)(a.counters + 9), 1ull);
      a.long_list[slot] = (int32_t)s0;
      return;
    }
    e0 = eL;
  }
  {
    double *t = target_cell(a, key);
    double s = *t;
    int64_t i = s0;
    if (e0 - i >= FB) {
      double cur[FB], nxt[FB];
#pragma unroll
      for (int k = 0; k < FB; k++) cur[k] = sv[i + k];
      for (; i + 2 * FB <= e0; i += FB) {
#pragma unroll
        for (int k = 0; k < FB; k++) nxt[k] = sv[i + FB + k];
#pragma unroll
        for (int k = 0; k < FB; k++) s += cur[k];
#pragma unroll
        for (int k = 0; k < FB; k++) cur[k] = nxt[k];
      }
#pragma unroll
      for (int k = 0; k < FB; k++) s += cur[k];
      i += FB;
    }
    for (; i < e0; i++) s += sv[i];
    *t = s;
  }
}

// s + v[0] + v[1] + ... + v[15], left to right, each addition rounded (IEEE
// v_add_f64, the same operation as `s += v`).  Written as one block of
// dependent additions so nothing is scheduled between them: the chain then
// runs at the latency of one addition per step while the next block's LDS
// reads are in flight.
__device__ inline double add16_in_order(double s, const double2 (&v)[8]) {
  asm volatile(
      "v_add_f64 %0, %0, %1\n\tv_add_f64 %0, %0, %2\n\t"
      "v_add_f64 %0, %0, %3\n\tv_add_f64 %0, %0, %4\n\t"
      "v_add_f64 %0, %0, %5\n\tv_add_f64 %0, %0, %6\n\t"
      "v_add_f64 %0, %0, %7\n\tv_add_f64 %0, %0, %8\n\t"
      "v_add_f64 %0, %0, %9\n\tv_add_f64 %0, %0, %10\n\t"
      "v_add_f64 %0, %0, %11\n\tv_add_f64 %0, %0, %12\n\t"
      "v_add_f64 %0, %0, %13\n\tv_add_f64 %0, %0, %14\n\t"
      "v_add_f64 %0, %0, %15\n\tv_add_f64 %0, %0, %16"
      : "+v"(s)
      : "v"(v[0].x), "v"(v[0].y), "v"(v[1].x), "v"(v[1].y), "v"(v[2].x), "v"(v[2].y), "v"(v[3].x), "v"(v[3].y),
        "v"(v[4].x), "v"(v[4].y), "v"(v[5].x), "v"(v[5].y), "v"(v[6].x), "v"(v[6].y), "v"(v[7].x), "v"(v[7].y));
  return s;
}

// One wavefront per long segment.  The wavefront loads the segment in
// coalesced chunks of LCH values (with their keys, which mark where the
// segment ends) into registers, one chunk ahead, and parks each chunk in LDS;
// lane 0 adds the parked values in order while the next chunk's loads are in
// flight.
constexpr int LU = 16;            // values per lane per chunk
constexpr int LCH = 64 * LU;      // values per chunk

__global__ __launch_bounds__(64) void fold_long_kernel(FoldArgs a) {
  const bool in1 = a.range && ranged_sort_in_k1(a.range);
  const uint32_t *__restrict__ sk = in1 ? a.sk1 : a.sk;
  const double *__restrict__ sv = in1 ? a.sv1 : a.sv;
  __shared__ __align__(16) double buf[LCH];
  const int lane = threadIdx.x;
  const int64_t nl = a.counters[9];
  for (int64_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const int64_t s0 = a.long_list[q];
    const uint32_t key = sk[s0];
    double *t = target_cell(a, key);
    double s = lane == 0 ? *t : 0.0;
    double v[LU];
    uint32_t kk[LU];   // keys of the chunk's positions (past the end: not the segment's)
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int64_t i = s0 + u * 64 + lane;
      kk[u] = i < a.n ? sk[i] : ~key;
      v[u] = i < a.n ? sv[i] : 0.0;
    }
    for (int64_t c0 = s0;; c0 += LCH) {
      // the segment's values in this chunk are a prefix of it, in position
      // order u * 64 + lane (keys are compared only here, after the previous
      // chunk's additions, so waiting for them never delays those)
      int m = 0;
#pragma unroll
      for (int u = 0; u < LU; u++) m += __popcll(__ballot(kk[u] == key));
      __syncthreads();   // lane 0 has finished with the previous chunk
#pragma unroll
      for (int u = 0; u < LU; u++) buf[u * 64 + lane] = v[u];
      __syncthreads();
      const bool more = m == LCH;
      if (more) {
        const int64_t c1 = c0 + LCH;
#pragma unroll
        for (int u = 0; u < LU; u++) {
          const int64_t i = c1 + u * 64 + lane;
          kk[u] = i < a.n ? sk[i] : ~key;
          v[u] = i < a.n ? sv[i] : 0.0;
        }
      }
      if (lane == 0) {
        // 32 parked values per round from two register sets: while one set's
        // 16 additions run, the other set is being refilled from LDS
        const double2 *b2 = reinterpret_cast<const double2 *>(buf);
        int i = 0;
        if (m >= 32) {
          double2 A[8], B[8];
#pragma unroll
          for (int k = 0; k < 8; k++) A[k] = b2[k];
#pragma unroll
          for (int k = 0; k < 8; k++) B[k] = b2[8 + k];
          // A holds values i..i+15, B holds i+16..i+31, i + 32 <= m
          for (; i + 64 <= m; i += 32) {
            s = add16_in_order(s, A);
#pragma unroll
            for (int k = 0; k < 8; k++) A[k] = b2[(i + 32) / 2 + k];
            s = add16_in_order(s, B);
#pragma unroll
            for (int k = 0; k < 8; k++) B[k] = b2[(i + 48) / 2 + k];
          }
          s = add16_in_order(s, A);
          s = add16_in_order(s, B);
          i += 32;
        }
        for (; i < m; i++) s += buf[i];
      }
      if (!more) break;
    }
    if (lane == 0) *t = s;
  }
}

}  // namespace

int deposit_cell_bits(int nz, int nr) {
  uint64_t cells = (uint64_t)(nz + 1) * (uint64_t)(nr + 1);
  int b = 1;
  while ((1ull << b) < cells) b++;
  return b;
}

// ordered scatter-add of the listed components into their target arrays
// ordered scatter-add of the listed components into their target arrays; with
// clear_first (at most three targets) the targets are set to +0 first
static hipError_t ordered_scatter(Run &R, const int *comps, int ncomp, double *const *targets, hipStream_t s,
                                  bool clear_first = false, bool zero_terms = false, double *sums = nullptr) {
  const int64_t n_ub = R.np_ub;
  if (clear_first && (n_ub == 0 || ncomp == 0))
    return launch_fill3(ncomp > 0 ? targets[0] : nullptr, ncomp > 1 ? targets[1] : nullptr,
                        ncomp > 2 ? targets[2] : nullptr, R.cells, 0.0, s);
  if (n_ub == 0 || ncomp == 0) return hipSuccess;
  const int64_t n = n_ub * ncomp;
  const int cell_bits = deposit_cell_bits(R.d.nz, R.d.nr);
  const int key_bits = cell_bits + 2;
  const uint32_t discard = (1u << key_bits) - 1u;   // all-ones cell with slot 3: never a real key
  KeyArgs k;
  k.nz = R.d.nz;
  k.nr = R.d.nr;
  k.ncomp = ncomp;
  k.ldz = R.ldz;
  for (int i = 0; i < 4; i++) k.comp[i] = i < ncomp ? comps[i] : 0;
  k.zmin = R.d.zmin;
  k.dz = R.d.dz;
  k.dr = R.d.dr;
  k.one_over_piDzDr2 = 1 / (3.141592653589793238462643 * R.d.dz * (R.d.dr * R.d.dr));
  k.Q = R.P[0];
  k.z = R.P[1];
  k.r = R.P[2];
  k.vz = R.P[6];
  k.vr = R.P[7];
  k.vphi = R.P[8];
  k.status = R.status;
  k.flags = R.flags;
  k.left_neu = R.d.left_bc == SPF_NEUMANN;
  k.right_neu = R.d.right_bc == SPF_NEUMANN;
  k.upper_neu = R.d.upper_bc == SPF_NEUMANN;
  k.np = R.counters;
  k.n_ub = n_ub;
  k.keys = R.keys;
  k.vals = R.cz;
  k.discard = discard;
  k.cell_bits = cell_bits;
  k.counters = R.counters;
  for (int i = 0; i < 3; i++) k.clear[i] = clear_first && i < ncomp ? targets[i] : nullptr;
  k.clear_n = clear_first ? R.cells : 0;
  k.zero_terms = zero_terms ? 1 : 0;
  k.zero_target = zero_terms ? targets[0] : nullptr;
  // enough workgroups for the particles and, when it clears the targets, for
  // a wide grid-stride clear (workgroups past n_ub only clear)
  int64_t kblocks = (n_ub + 255) / 256;
  if (clear_first) kblocks = std::max<int64_t>(kblocks, std::min<int64_t>(8192, (R.cells + 255) / 256));
  // Over keys wider than three 7-bit digits the sort takes the key range
  // found by the key kernel (a bunch spans a band of rows, so it often needs a
  // pass fewer); narrower keys keep the plain sort, whose three passes the
  // range could not shorten and whose set-up it would only add to.
  const bool ranged = key_bits > 21;
  k.krange = ranged ? reinterpret_cast<unsigned long long *>(R.krange) : nullptr;
  if (ranged) hipLaunchKernelGGL(key_range_reset_kernel, dim3(1), dim3(kRangeSlots), 0, s, k.krange);
  hipLaunchKernelGGL(key_kernel, dim3((unsigned)kblocks), dim3(256), 0, s, k);
  if (ranged) hipLaunchKernelGGL(key_range_final_kernel, dim3(1), dim3(kRangeSlots), 0, s, k.krange);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the first component's values in particle order, before the sort reuses their buffer
  if (sums && (e = launch_sum2_vec(R.cz, n_ub, R.red, R.red_n, sums, s)) != hipSuccess) return e;
  const uint64_t *range = ranged ? reinterpret_cast<const uint64_t *>(R.krange + 2 * kRangeSlots) : nullptr;
  uint32_t *sk = R.keys;
  double *sv = R.cz;
  e = ranged ? radix_sort_pairs_ranged(R.keys, R.cz, R.skeys, R.scz, n, key_bits, range, discard, R.sort_tmp,
                                       R.sort_tmp_bytes, s)
             : radix_sort_pairs(R.keys, R.cz, R.skeys, R.scz, n, key_bits, R.sort_tmp, R.sort_tmp_bytes, &sk, &sv, s);
  if (e != hipSuccess) return e;
  FoldArgs f;
  f.nz = R.d.nz;
  f.ldz = R.ldz;
  f.cell_bits = cell_bits;
  f.n = n;
  f.discard = discard;
  f.sk = sk;
  f.sv = sv;
  f.sk1 = R.skeys;
  f.sv1 = R.scz;
  f.range = range;
  f.counters = R.counters;
  f.long_list = R.long_list;
  for (int i = 0; i < 4; i++) f.target[i] = i < ncomp ? targets[i] : nullptr;
  hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // at most n / LONG_SEG long segments
  int64_t nlong = n / LONG_SEG + 1;
  hipLaunchKernelGGL(fold_long_kernel, dim3((unsigned)std::min<int64_t>(nlong, 8192)), dim3(64), 0, s, f);
  return hipGetLastError();
}

hipError_t deposit_currents(Run &R, hipStream_t s, bool clear_first) {
  int comps[3];
  double *targets[3];
  int nc = 0;
  if (R.d.modes & SPF_TM_FIELDS) {
    comps[nc] = COMP_JZ;
    targets[nc++] = R.Jz;
    comps[nc] = COMP_JR;
    targets[nc++] = R.Jr;
  }
  if (R.d.modes & SPF_TE_FIELDS) {
    comps[nc] = COMP_JPHI;
    targets[nc++] = R.Jphi;
  }
  return ordered_scatter(R, comps, nc, targets, s, clear_first);
}

hipError_t launch_q_subtract(Run &R, hipStream_t s, double *qreal_sums) {
  int comps[1] = {COMP_Q};
  double *targets[1] = {R.Q};
  if (qreal_sums && R.np_ub == 0) return launch_fill3(qreal_sums, nullptr, nullptr, 2, 0.0, s);
  return ordered_scatter(R, comps, 1, targets, s, false, true, qreal_sums);
}

namespace {

// Qreal of src/poisson_correction.c:221-224 (no smearing: fb = fl = 1) summed in
// particle order from +0, one rounded addition at a time; the zero-weight
// terms of :226-239 add +-0, which never changes a sum that starts at +0.
__global__ __launch_bounds__(64) void qreal_serial_kernel(KeyArgs a, double *out) {
  __shared__ double buf[64];
  const int lane = threadIdx.x;
  const int64_t np = *a.np;
  double s = 0.0;
  for (int64_t c = 0; c < np; c += 64) {
    const int64_t ip = c + lane;
    double t = 0.0;
    bool take = false;
    if (ip < np && (a.status[ip] & SPF_PART_ACTIVE)) {
      const double th = (a.z[ip] - a.zmin) / a.dz + 0.5, yh = a.r[ip] / a.dr + 0.5;
      if (th > -1.0 && th < a.nz + 1.0 && yh > -1.0 && yh < a.nr + 1.0) {
        const long long iz = (long long)th, ir = (long long)yh;
        if (iz >= 0 && iz < a.nz && ir >= 0 && ir < a.nr) {
          const int16_t f = a.flags[fix(a.ldz, (int)iz, (int)ir)];
          take = !((f & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (iz == 0 && a.left_neu) ||
                   (iz >= a.nz - 1 && a.right_neu) || (ir >= a.nr - 1 && a.upper_neu));
          t = a.Q[ip] * 1.0 * 1.0;
        }
      }
    }
    buf[lane] = take ? t : 0.0;
    const unsigned long long m = __ballot(take);
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 64; k++)
        if ((m >> k) & 1ull) s += buf[k];
    __syncthreads();
  }
  if (lane == 0) *out = s;
}

}  // namespace

hipError_t launch_qreal_serial(Run &R, double *out, hipStream_t s) {
  KeyArgs k{};
  k.nz = R.d.nz;
  k.nr = R.d.nr;
  k.ldz = R.ldz;
  k.zmin = R.d.zmin;
  k.dz = R.d.dz;
  k.dr = R.d.dr;
  k.Q = R.P[0];
  k.z = R.P[1];
  k.r = R.P[2];
  k.status = R.status;
  k.flags = R.flags;
  k.left_neu = R.d.left_bc == SPF_NEUMANN;
  k.right_neu = R.d.right_bc == SPF_NEUMANN;
  k.upper_neu = R.d.upper_bc == SPF_NEUMANN;
  k.np = R.counters;
  hipLaunchKernelGGL(qreal_serial_kernel, dim3(1), dim3(64), 0, s, k, out);
  return hipGetLastError();
}

}  // namespace spf
