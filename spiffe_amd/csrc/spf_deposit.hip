// spf_deposit.hip -- nearest-grid-point current deposition
// (src/space_charge.c:34-295) and the particle-charge removal of the Poisson
// correction (src/poisson_correction.c:184-243, :377-437), with the
// reference's per-cell summation order and no floating-point atomics.
//
// The reference adds each particle's contribution to its cell in particle
// index order, three passes (Jz, Jr, Jphi).  Floating-point addition is not
// associative, so the device keeps that order exactly:
//  1. key: each active particle gets the cell (iz_h, ir_h) of the Jphi map
//     (iz = (z-zmin)/dz + 0.5, ir = r/dr + 0.5, C truncation); the Jz map
//     differs only by iz_a = (long)(x+1.0) in {iz_h, iz_h+1} and the Jr map by
//     ir_a = (long)(y+1.0) in {ir_h, ir_h+1}, recorded as two bits;
//  2. a stable radix sort by key keeps index order inside every group;
//  3. one thread per non-empty group folds its cells sequentially: Jphi from
//     its own group, Jz by merging (by particle index) the group with its
//     z-neighbour group, Jr with its r-neighbour group.  Every J cell has
//     exactly one writer.
// Contributions that are exactly zero in the reference (the ir >= nr-1 and
// Jr ir == 0 branches) are skipped: J starts each step at +0 and a running
// sum of nonzero doubles from +0 is never -0, so adding a zero cannot change
// a bit.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "spf_internal.h"

namespace spf {

namespace {

constexpr uint32_t DISCARD = 0xffffffffu;
constexpr uint8_t B_JZ = 1, B_DZ = 2, B_JR = 4, B_DR = 8, B_JPHI = 16;

struct KeyArgs {
  int nz, nr, tm, te;
  double zmin, dz, dr, one_over_piDzDr2;
  const double *Q, *z, *r, *vz, *vr, *vphi;
  const uint32_t *status;
  const int64_t *np;
  uint32_t *keys;
  int32_t *vals;
  uint8_t *bits;
  double *cz, *cr, *cph;
  uint32_t discard;
};

// NGP value of the Jz/Jphi loops for cell row ir (0 <= ir <= nr-2), :92-100, :254-262
__device__ inline double jm_centered(double QVG, long long ir) {
  if (ir == 0) return 4.0 * 1.0 * QVG;
  return 1.0 * QVG / ((ir + 0.5) * (ir + 0.5) - (ir - 0.5) * (ir - 0.5));
}

__global__ __launch_bounds__(256) void key_kernel(KeyArgs a, int64_t n_ub) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n_ub) return;
  int64_t np = *a.np;
  a.vals[ip] = (int32_t)ip;
  if (ip >= np || !(a.status[ip] & SPF_PART_ACTIVE)) {
    a.keys[ip] = a.discard;
    return;
  }
  double x = (a.z[ip] - a.zmin) / a.dz;
  double y = a.r[ip] / a.dr;
  double th = x + 0.5, t1 = x + 1.0, yh = y + 0.5, y1 = y + 1.0;
  // iz_h in [-1, nz] and ir_h in [-1, nr] are the only groups any component
  // can use; the bounds also keep the conversions in range (NaN fails them)
  if (!(th > -2.0 && th < a.nz + 1.0 && yh > -2.0 && yh < a.nr + 1.0)) {
    a.keys[ip] = a.discard;
    return;
  }
  long long izh = (long long)th, izA = (long long)t1, irh = (long long)yh, irA = (long long)y1;
  int nz = a.nz, nr = a.nr;
  double Qo = a.Q[ip] * a.one_over_piDzDr2;
  uint8_t b = 0;
  if (izA - izh == 1) b |= B_DZ;
  if (irA - irh == 1) b |= B_DR;
  double cz = 0, cr = 0, cph = 0;
  if (a.tm) {
    // Jz: iz = x+1.0, ir = y+0.5, valid iz in [0,nz], ir in [0,nr], nonzero for ir <= nr-2
    if (izA >= 0 && izA <= nz && irh >= 0 && irh <= nr - 2) {
      cz = jm_centered(a.vz[ip] * Qo, irh);
      b |= B_JZ;
    }
    // Jr: iz = x+0.5, ir = y+1.0, valid iz in [0,nz], ir in [0,nr-1], nonzero for 1 <= ir <= nr-2
    if (izh >= 0 && izh <= nz && irA >= 1 && irA <= nr - 2) {
      double QVG = a.vr[ip] * Qo;
      cr = 1.0 * QVG / ((double)irA * (double)irA - (double)(irA - 1) * (double)(irA - 1));
      b |= B_JR;
    }
  }
  if (a.te) {
    if (izh >= 0 && izh <= nz && irh >= 0 && irh <= nr - 2) {
      cph = jm_centered(a.vphi[ip] * Qo, irh);
      b |= B_JPHI;
    }
  }
  a.keys[ip] = (uint32_t)((irh + 1) * (nz + 2) + (izh + 1));
  a.bits[ip] = b;
  a.cz[ip] = cz;
  a.cr[ip] = cr;
  a.cph[ip] = cph;
}

__global__ void init_groups_kernel(int64_t *counters, int64_t n_ub) {
  counters[5] = 0;     // G
  counters[6] = n_ub;  // nvalid
}

__global__ __launch_bounds__(256) void mark_kernel(const uint32_t *sk, int64_t n, uint8_t *gflag, int64_t *counters,
                                                   uint32_t discard) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = sk[i];
  uint32_t prev = i ? sk[i - 1] : 0xfffffffeu;
  gflag[i] = (k != discard && (i == 0 || prev != k)) ? 1 : 0;
  if (k == discard && (i == 0 || prev != discard)) counters[6] = i;
}

__global__ __launch_bounds__(256) void permute_kernel(const uint32_t *sk, const int32_t *sv, int64_t n,
                                                      const int64_t *counters, const double *cz, const double *cr,
                                                      const double *cph, const uint8_t *bits, const double *Q,
                                                      double *scz, double *scr, double *scph, uint8_t *sbits,
                                                      double *sq) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= counters[6]) return;
  int32_t ip = sv[i];
  scz[i] = cz[ip];
  scr[i] = cr[ip];
  scph[i] = cph[ip];
  sbits[i] = bits[ip];
  sq[i] = Q[ip];
}

struct FoldArgs {
  int nz, nr, ldz, tm, te;
  const uint32_t *sk;
  const int32_t *sv;
  const int32_t *gstart;
  const int64_t *counters;
  const double *scz, *scr, *scph, *sq;
  const uint8_t *sbits;
  double *Jz, *Jr, *Jphi;
  // Poisson charge removal
  double *Q;
  const int16_t *flags;
  int left_neu, right_neu, upper_neu;
  double *gq;   // per-group removed charge (for Qreal), may be null
};

__device__ inline uint32_t gkey(const FoldArgs &a, int64_t g) { return a.sk[a.gstart[g]]; }

__device__ int64_t find_group(const FoldArgs &a, int64_t G, uint32_t key) {
  int64_t lo = 0, hi = G - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    uint32_t k = gkey(a, mid);
    if (k == key) return mid;
    if (k < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

__device__ inline void grange(const FoldArgs &a, int64_t G, int64_t g, int64_t &s, int64_t &e) {
  s = a.gstart[g];
  e = (g + 1 < G) ? a.gstart[g + 1] : a.counters[6];
}

// sequential fold of entries selected from two index-sorted ranges, merged by particle index
template <typename SelA, typename SelB>
__device__ double merge_fold(const FoldArgs &a, double s, const double *val, int64_t a0, int64_t a1, SelA selA,
                             int64_t b0, int64_t b1, SelB selB) {
  int64_t i = a0, j = b0;
  while (true) {
    while (i < a1 && !selA(a.sbits[i])) i++;
    while (j < b1 && !selB(a.sbits[j])) j++;
    if (i >= a1 && j >= b1) break;
    bool takeA = (j >= b1) || (i < a1 && a.sv[i] < a.sv[j]);
    if (takeA) {
      s += val[i];
      i++;
    } else {
      s += val[j];
      j++;
    }
  }
  return s;
}

__global__ __launch_bounds__(256) void fold_kernel(FoldArgs a) {
  const int64_t G = a.counters[5];
  const int NZG = a.nz + 2;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x) {
    int64_t s0, e0;
    grange(a, G, g, s0, e0);
    uint32_t key = a.sk[s0];
    int gz = (int)(key % NZG), gr = (int)(key / NZG);
    int izh = gz - 1, irh = gr - 1;
    // Jphi(iz_h, ir_h), :223-293
    if (a.te && izh >= 0 && izh <= a.nz && irh >= 0 && irh <= a.nr - 2) {
      int64_t p = fix(a.ldz, izh, irh);
      double s = a.Jphi[p];
      for (int64_t i = s0; i < e0; i++)
        if (a.sbits[i] & B_JPHI) s += a.scph[i];
      a.Jphi[p] = s;
    }
    if (!a.tm) continue;
    // Jz, :61-139: cell (iz_h, ir_h) gets this group's dz=0 entries merged with
    // the (iz_h-1, ir_h) group's dz=1 entries
    if (irh >= 0 && irh <= a.nr - 2) {
      if (izh >= 0 && izh <= a.nz) {
        int64_t ps = 0, pe = 0;
        if (gz > 0 && g > 0 && gkey(a, g - 1) == key - 1) grange(a, G, g - 1, ps, pe);
        int64_t p = fix(a.ldz, izh, irh);
        a.Jz[p] = merge_fold(
            a, a.Jz[p], a.scz, s0, e0, [](uint8_t b) { return (b & B_JZ) && !(b & B_DZ); }, ps, pe,
            [](uint8_t b) { return (b & B_JZ) && (b & B_DZ); });
      }
      // cell (iz_h+1, ir_h) when no (iz_h+1, ir_h) group exists to own it
      bool next_exists = (g + 1 < G) && gz + 1 < NZG && gkey(a, g + 1) == key + 1;
      if (!next_exists && izh + 1 >= 0 && izh + 1 <= a.nz) {
        int64_t p = fix(a.ldz, izh + 1, irh);
        a.Jz[p] = merge_fold(
            a, a.Jz[p], a.scz, s0, e0, [](uint8_t b) { return (b & B_JZ) && (b & B_DZ); }, 0, 0,
            [](uint8_t) { return false; });
      }
    }
    // Jr, :141-215: cell (iz_h, ir_h) gets this group's dr=0 entries merged
    // with the (iz_h, ir_h-1) group's dr=1 entries
    if (izh >= 0 && izh <= a.nz) {
      if (irh >= 1 && irh <= a.nr - 2) {
        int64_t ps = 0, pe = 0;
        if (gr > 0) {
          int64_t q = find_group(a, G, key - (uint32_t)NZG);
          if (q >= 0) grange(a, G, q, ps, pe);
        }
        int64_t p = fix(a.ldz, izh, irh);
        a.Jr[p] = merge_fold(
            a, a.Jr[p], a.scr, s0, e0, [](uint8_t b) { return (b & B_JR) && !(b & B_DR); }, ps, pe,
            [](uint8_t b) { return (b & B_JR) && (b & B_DR); });
      }
      int irn = irh + 1;
      if (irn >= 1 && irn <= a.nr - 2) {
        bool up_exists = gr + 1 <= a.nr + 1 && find_group(a, G, key + (uint32_t)NZG) >= 0;
        if (!up_exists) {
          int64_t p = fix(a.ldz, izh, irn);
          a.Jr[p] = merge_fold(
              a, a.Jr[p], a.scr, s0, e0, [](uint8_t b) { return (b & B_JR) && (b & B_DR); }, 0, 0,
              [](uint8_t) { return false; });
        }
      }
    }
  }
}

// Q[iz][ir] -= Q_p for the particles of group (iz, ir), in index order,
// src/poisson_correction.c:190-242 with fb = fl = 1 (no smearing).
__global__ __launch_bounds__(256) void q_fold_kernel(FoldArgs a) {
  const int64_t G = a.counters[5];
  const int NZG = a.nz + 2;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < G; g += (int64_t)gridDim.x * blockDim.x) {
    int64_t s0, e0;
    grange(a, G, g, s0, e0);
    uint32_t key = a.sk[s0];
    int iz = (int)(key % NZG) - 1, ir = (int)(key / NZG) - 1;
    double removed = 0;
    if (iz >= 0 && iz < a.nz && ir >= 0 && ir < a.nr) {
      int64_t p = fix(a.ldz, iz, ir);
      bool blocked = (a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (iz == 0 && a.left_neu) ||
                     (iz >= a.nz - 1 && a.right_neu) || (ir >= a.nr - 1 && a.upper_neu);
      if (!blocked) {
        double s = a.Q[p];
        for (int64_t i = s0; i < e0; i++) {
          double c = a.sq[i] * 1.0 * 1.0;
          s -= c;
          removed += c;
        }
        a.Q[p] = s;
      }
    }
    if (a.gq) a.gq[g] = removed;
  }
}

FoldArgs fold_args(Run &R) {
  FoldArgs f;
  f.nz = R.d.nz;
  f.nr = R.d.nr;
  f.ldz = R.ldz;
  f.tm = (R.d.modes & SPF_TM_FIELDS) != 0;
  f.te = (R.d.modes & SPF_TE_FIELDS) != 0;
  f.sk = R.skeys;
  f.sv = R.svals;
  f.gstart = R.gstart;
  f.counters = R.counters;
  f.scz = R.scz;
  f.scr = R.scr;
  f.scph = R.scph;
  f.sq = R.sq;
  f.sbits = R.sbits;
  f.Jz = R.Jz;
  f.Jr = R.Jr;
  f.Jphi = R.Jphi;
  f.Q = R.Q;
  f.flags = R.flags;
  f.left_neu = R.d.left_bc == SPF_NEUMANN;
  f.right_neu = R.d.right_bc == SPF_NEUMANN;
  f.upper_neu = R.d.upper_bc == SPF_NEUMANN;
  f.gq = nullptr;
  return f;
}

unsigned fold_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (unsigned)b;
}

}  // namespace

size_t deposit_sort_bytes(int64_t n, int key_bits) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, (int32_t *)nullptr,
                            (int32_t *)nullptr, (size_t)n, 0u, (unsigned)key_bits, (hipStream_t)0);
  return bytes;
}

hipError_t deposit_prepare(Run &R, hipStream_t s) {
  int64_t n = R.np_ub;
  hipLaunchKernelGGL(init_groups_kernel, dim3(1), dim3(1), 0, s, R.counters, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n == 0) {
    R.sort_valid = (e == hipSuccess);
    return e;
  }
  KeyArgs k;
  k.nz = R.d.nz;
  k.nr = R.d.nr;
  k.tm = (R.d.modes & SPF_TM_FIELDS) != 0;
  k.te = (R.d.modes & SPF_TE_FIELDS) != 0;
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
  k.np = R.counters;
  k.keys = R.keys;
  k.vals = R.vals;
  k.bits = R.dbits;
  k.cz = R.cz;
  k.cr = R.cr;
  k.cph = R.cph;
  k.discard = (R.key_bits >= 32) ? DISCARD : ((1u << R.key_bits) - 1u);
  hipLaunchKernelGGL(key_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, k, n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t bytes = R.sort_tmp_bytes;
  e = rocprim::radix_sort_pairs(R.sort_tmp, bytes, R.keys, R.skeys, R.vals, R.svals, (size_t)n, 0u,
                                (unsigned)R.key_bits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mark_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.skeys, n, R.gflag,
                     R.counters, k.discard);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = ordered_compact(R.gflag, n, R.gstart, R.counters + 5, R.bcount, R.boff, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.skeys, R.svals, n,
                     R.counters, R.cz, R.cr, R.cph, R.dbits, R.P[0], R.scz, R.scr, R.scph, R.sbits, R.sq);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  R.sort_valid = true;
  return hipSuccess;
}

hipError_t deposit_fold(Run &R, hipStream_t s) {
  if (R.np_ub == 0) return hipSuccess;
  FoldArgs f = fold_args(R);
  hipLaunchKernelGGL(fold_kernel, dim3(fold_grid(R.np_ub)), dim3(256), 0, s, f);
  return hipGetLastError();
}

hipError_t launch_q_subtract(Run &R, hipStream_t s) {
  if (R.np_ub == 0) return hipSuccess;
  FoldArgs f = fold_args(R);
  hipLaunchKernelGGL(q_fold_kernel, dim3(fold_grid(R.np_ub)), dim3(256), 0, s, f);
  return hipGetLastError();
}

}  // namespace spf
