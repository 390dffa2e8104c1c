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
  int need_v[3];                // vz, vr, vphi are read (a component uses them)
  double *pvals;                // reduce path: first component's value per particle (or null)
  int32_t *wg_count;            // reduce path: partials per workgroup
  int smear_r, smear_z;         // FL_RADIAL_SMEAR, FL_LONGIT_SMEAR
  int tpp;                      // key slots per particle and component: 4 when smeared, else 1
  double *jzero[3];             // J targets of slot k that may hold -0 (not cleared first), or null
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

// the columns of one particle slot the deposition reads, loaded together so
// that no load waits for another (slots past n_ub load nothing)
struct PartVals {
  bool live;
  double Q, z, r, vz, vr, vphi;
};

__device__ inline PartVals load_part(const KeyArgs &a, int64_t ip, int64_t np) {
  PartVals p{false, 0, 0, 0, 0, 0, 0};
  if (ip < a.n_ub) {
    const uint32_t st = a.status[ip];
    p.Q = a.Q[ip];
    p.z = a.z[ip];
    p.r = a.r[ip];
    if (a.need_v[0]) p.vz = a.vz[ip];
    if (a.need_v[1]) p.vr = a.vr[ip];
    if (a.need_v[2]) p.vphi = a.vphi[ip];
    p.live = ip < np && (st & SPF_PART_ACTIVE);
  }
  return p;
}

// cell indices of one particle that the components' keys are made from
struct PartCells {
  bool live;
  long long izh, izA, irh, irA;   // x+0.5, x+1.0, y+0.5, y+1.0 truncated (-1: outside)
  double Qo;
};

__device__ inline PartCells part_cells(const KeyArgs &a, const PartVals &pv) {
  PartCells c{pv.live, -1, -1, -1, -1, 0.0};
  if (pv.live) {
    double x = (pv.z - a.zmin) / a.dz;
    double y = pv.r / a.dr;
    double th = x + 0.5, t1 = x + 1.0, yh = y + 0.5, y1 = y + 1.0;
    // every valid cell index lies in [0, nz] x [0, nr]; outside these
    // bounds (and for NaN) the conversions are skipped
    if (th > -1.0 && th < a.nz + 1.0) c.izh = (long long)th;
    if (t1 > -1.0 && t1 < a.nz + 1.0) c.izA = (long long)t1;
    if (yh > -1.0 && yh < a.nr + 1.0) c.irh = (long long)yh;
    if (y1 > -1.0 && y1 < a.nr + 1.0) c.irA = (long long)y1;
    c.Qo = pv.Q * a.one_over_piDzDr2;
  }
  return c;
}

// the key (or `discard`) and value of component slot k of particle pv
__device__ inline uint32_t part_contrib(const KeyArgs &a, const PartCells &pc, const PartVals &pv, int k, double &v) {
  const int c = a.comp[k];
  const int nz = a.nz, nr = a.nr;
  long long iz = -1, ir = -1;
  v = 0;
  if (pc.live) {
    const long long izh = pc.izh, izA = pc.izA, irh = pc.irh, irA = pc.irA;
    const double Qo = pc.Qo;
    if (c == COMP_JZ) {
      // iz = x+1.0, ir = y+0.5; valid iz <= nz, ir <= nr; nonzero only for ir <= nr-2
      if (izA >= 0 && izA <= nz && irh >= 0 && irh <= nr - 2) {
        iz = izA;
        ir = irh;
        v = jm_centered(pv.vz * Qo, irh);
      }
    } else if (c == COMP_JR) {
      // iz = x+0.5, ir = y+1.0; valid ir <= nr-1; nonzero only for 1 <= ir <= nr-2
      if (izh >= 0 && izh <= nz && irA >= 1 && irA <= nr - 2) {
        iz = izh;
        ir = irA;
        double QVG = pv.vr * Qo;
        v = 1.0 * QVG / ((double)irA * (double)irA - (double)(irA - 1) * (double)(irA - 1));
      }
    } else if (c == COMP_JPHI) {
      if (izh >= 0 && izh <= nz && irh >= 0 && irh <= nr - 2) {
        iz = izh;
        ir = irh;
        v = jm_centered(pv.vphi * Qo, irh);
      }
    } else {  // COMP_Q: iz = x+0.5, ir = y+0.5 in [0,nz-1] x [0,nr-1], not metal or a Neumann edge
      if (izh >= 0 && izh < nz && irh >= 0 && irh < nr) {
        int16_t f = a.flags[fix(a.ldz, (int)izh, (int)irh)];
        bool blocked = (f & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (izh == 0 && a.left_neu) ||
                       (izh >= nz - 1 && a.right_neu) || (irh >= nr - 1 && a.upper_neu);
        if (!blocked) {
          iz = izh;
          ir = irh;
          v = -(pv.Q * 1.0 * 1.0);
        }
        // The reference also subtracts Q*fb*fr, Q*fa*fl and Q*fa*fr from the
        // cells (iz+1, ir), (iz, ir+1), (iz+1, ir+1) (src/poisson_correction.c:
        // 222-236); without smearing these are Q*0, a zero with the sign of Q.
        // Subtracting +0 changes no value; subtracting -0 (Q < 0) turns a -0
        // cell into +0 and leaves every other value alone.  A cell can hold -0
        // only while nothing but -0 has been added to a -0 start, so turning
        // -0 into +0 before the ordered fold gives the reference's bits
        // whatever the position of the zero terms in its sequence.
        if (a.zero_terms && __builtin_signbit(pv.Q)) {
          const long long zc[3][2] = {{izh + 1, irh}, {izh, irh + 1}, {izh + 1, irh + 1}};
          const bool zb[3] = {(izh == 0 && a.left_neu) || (izh >= nz && a.right_neu) || (irh >= nr - 1 && a.upper_neu),
                              (izh == 0 && a.left_neu) || (izh >= nz - 1 && a.right_neu) || (irh >= nr && a.upper_neu),
                              (izh == 0 && a.left_neu) || (izh >= nz && a.right_neu) || (irh >= nr && a.upper_neu)};
#pragma unroll
          for (int t = 0; t < 3; t++) {
            const int64_t pcell = fix(a.ldz, (int)zc[t][0], (int)zc[t][1]);
            if (zb[t] || (a.flags[pcell] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) continue;
            double *q = a.zero_target + pcell;
            if (__double_as_longlong(*q) == (long long)0x8000000000000000ull) *q = 0.0;
          }
        }
      }
    }
  }
  if (iz < 0) return a.discard;
  const uint32_t cell = (uint32_t)(ir * (nz + 1) + iz);
  return (cell << 2) | (uint32_t)k;
}

// ---- area-weighted ("smeared") deposition, src/space_charge.c:61-293 and
// src/poisson_correction.c:188-239 with FL_RADIAL_SMEAR / FL_LONGIT_SMEAR.
// A particle adds to up to four points per component, in the reference's
// order (iz, ir), (iz+1, ir), (iz, ir+1), (iz+1, ir+1).  Its key slots are
// laid out particle-major (slot 4*ip + t of the component's block), so the
// stable sort keeps each point's terms in particle index order.  Every term
// the reference adds inside the arrays is kept, zeros included, so signed
// zeros come out as the reference's.  The one term the reference writes past
// its arrays (row nz+1, from a particle with iz == nz) is left out.

// C's conversion to long of v when -1 < v < lim (the conversion truncates
// toward zero), else -1: every index the callers accept lies below lim
__device__ inline long long trunc_in(double v, double lim) { return (v > -1.0 && v < lim) ? (long long)v : -1; }

struct Smear4 {
  long long iz[4], ir[4];   // iz[t] = -1: no term
  double v[4];
};

// component c (COMP_JZ, COMP_JR, COMP_JPHI) of particle pv, src/space_charge.c:61-139, :141-215, :222-293
__device__ inline void smear_current(const KeyArgs &a, const PartVals &pv, int c, Smear4 &s) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    s.iz[t] = -1;
    s.ir[t] = 0;
    s.v[t] = 0.0;
  }
  if (!pv.live) return;
  const bool RS = a.smear_r, LS = a.smear_z;
  const long long nz = a.nz, nr = a.nr;
  const double x = (pv.z - a.zmin) / a.dz, y = pv.r / a.dr;
  long long iz, ir;
  double vel;
  if (c == COMP_JZ) {
    ir = trunc_in(RS ? y : y + 0.5, nr + 1.0);
    iz = trunc_in(LS ? x + 0.5 : x + 1.0, nz + 1.0);
    vel = pv.vz;
  } else if (c == COMP_JR) {
    ir = trunc_in(RS ? y + 0.5 : y + 1.0, nr + 1.0);
    iz = trunc_in(LS ? x : x + 0.5, nz + 1.0);
    vel = pv.vr;
  } else {
    ir = trunc_in(RS ? y : y + 0.5, nr + 1.0);
    iz = trunc_in(LS ? x : x + 0.5, nz + 1.0);
    vel = pv.vphi;
  }
  if (iz < 0 || iz > nz || ir < 0 || ir > (c == COMP_JR ? nr - 1 : nr)) return;
  const double QVG = vel * (pv.Q * a.one_over_piDzDr2);
  // radial weights and the current densities of the rows below and above
  double fp, fm, Jm, Jp;
  if (c == COMP_JR) {
    if (RS) {
      fp = (y + 0.5) - (double)ir;
      fm = 1 - fp;
    } else {
      fm = 1;
      fp = 0;
    }
    if (ir == 0) {
      Jm = 0;
      Jp = QVG;
    } else if (ir < nr - 1) {
      Jm = fm * QVG / ((double)ir * (double)ir - (double)(ir - 1) * (double)(ir - 1));
      Jp = fp * QVG / ((double)(ir + 1) * (double)(ir + 1) - (double)ir * (double)ir);
    } else {
      Jm = 0;
      ir--;
      Jp = QVG / ((double)(ir + 1) * (double)(ir + 1) - (double)ir * (double)ir);
    }
  } else {
    if (RS) {
      fp = y - (double)ir;
      fm = 1 - fp;
    } else {
      fm = 1;
      fp = 0;
    }
    if (ir == 0) {
      Jm = 4 * fm * QVG;
      Jp = fp * QVG / (1.5 * 1.5 - 0.5 * 0.5);
    } else if (ir < nr - 1) {
      Jm = fm * QVG / ((ir + 0.5) * (ir + 0.5) - (ir - 0.5) * (ir - 0.5));
      Jp = fp * QVG / ((ir + 1.5) * (ir + 1.5) - (ir + 0.5) * (ir + 0.5));
    } else {
      Jm = 0;
      Jp = fp * QVG / ((ir + 0.5) * (ir + 0.5) - (ir - 0.5) * (ir - 0.5));
      ir--;
    }
  }
  // longitudinal weights of the points to the left and right
  double gm = 1, gp = 0;
  if (LS) {
    if (c == COMP_JZ) {
      if (iz == 0) {
        gp = 1;
        gm = 0;
      } else if (iz == nz) {
        gm = 1;
        gp = 0;
      } else {
        gp = (x + 0.5) - (double)iz;
        gm = 1 - gp;
      }
    } else if (c == COMP_JR) {
      if (iz == nz) {
        gm = 1;
        gp = 0;
      } else {
        gp = x - (double)iz;
        gm = 1 - gp;
      }
    } else {
      if (iz == nz - 1) {
        gp = 0;
        gm = 1 - gp;
      } else {
        gp = x - (double)iz;
        gm = 1 - gp;
      }
    }
  }
  s.iz[0] = iz;
  s.ir[0] = ir;
  s.v[0] = gm * Jm;
  if (LS && iz + 1 <= nz) {
    s.iz[1] = iz + 1;
    s.ir[1] = ir;
    s.v[1] = gp * Jm;
  }
  if (RS) {
    s.iz[2] = iz;
    s.ir[2] = ir + 1;
    s.v[2] = gm * Jp;
    if (LS && iz + 1 <= nz) {
      s.iz[3] = iz + 1;
      s.ir[3] = ir + 1;
      s.v[3] = gp * Jp;
    }
  }
}

// the charge a particle takes off each of its four points, src/poisson_correction.c:188-239:
// v[t] = Q*weight (the reference subtracts it and adds it to Qreal); iz[t] = -1
// where the point is metal, on a Neumann edge, or the particle outside
__device__ inline void smear_charge(const KeyArgs &a, const PartVals &pv, Smear4 &s) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    s.iz[t] = -1;
    s.ir[t] = 0;
    s.v[t] = 0.0;
  }
  if (!pv.live) return;
  const bool RS = a.smear_r, LS = a.smear_z;
  const long long nz = a.nz, nr = a.nr;
  const double x = (pv.z - a.zmin) / a.dz, y = pv.r / a.dr;
  const long long iz = trunc_in(LS ? x : x + 0.5, nz + 1.0), ir = trunc_in(RS ? y : y + 0.5, nr + 1.0);
  if (iz < 0 || iz >= nz || ir < 0 || ir >= nr) return;
  double fa, fb, fl, fr;
  if (RS && ir != nr - 1) {
    fa = y - (double)ir;
    fb = 1 - fa;
  } else {
    fb = 1;
    fa = 0;
  }
  if (LS && iz != nz - 1) {
    fr = x - (double)iz;
    fl = 1 - fr;
  } else {
    fl = 1;
    fr = 0;
  }
  const bool L = a.left_neu, R = a.right_neu, U = a.upper_neu;
  const long long pz[4] = {iz, iz + 1, iz, iz + 1}, pr[4] = {ir, ir, ir + 1, ir + 1};
  const bool edge[4] = {(iz == 0 && L) || (iz >= nz - 1 && R) || (ir >= nr - 1 && U),
                        (iz == 0 && L) || (iz >= nz && R) || (ir >= nr - 1 && U),
                        (iz == 0 && L) || (iz >= nz - 1 && R) || (ir >= nr && U),
                        (iz == 0 && L) || (iz >= nz && R) || (ir >= nr && U)};
  const double w[4] = {pv.Q * fb * fl, pv.Q * fb * fr, pv.Q * fa * fl, pv.Q * fa * fr};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int16_t f = a.flags[fix(a.ldz, (int)pz[t], (int)pr[t])];
    if (edge[t] || (f & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) continue;
    s.iz[t] = pz[t];
    s.ir[t] = pr[t];
    s.v[t] = w[t];
  }
}

// The reference's nearest-point loops also add +0 (a literal zero times a
// weight of 1) to J at points this deposition leaves out: for Jz and Jphi the
// row below when ir >= nr-1 (:92-101, :254-263), for Jr row 0 when ir == 0
// and row nr-2 when ir == nr-1 (:183-193).  Adding +0 changes only a -0 (into
// +0), and a point holds -0 only while nothing but -0 has been added to a -0
// start, so turning such a -0 into +0 before the ordered fold gives the
// reference's bits wherever the zero falls in its sequence.  J is +0 after a
// clear, so this matters only when J is not cleared first (the J time
// filter, or a deposition onto uploaded currents).
__device__ inline void plus_zero(double *T, int ldz, long long iz, long long ir) {
  double *q = T + fix(ldz, (int)iz, (int)ir);
  if (__double_as_longlong(*q) == (long long)0x8000000000000000ull) *q = 0.0;
}

__device__ inline void dropped_j_zeros(const KeyArgs &a, const PartCells &pc) {
  if (!pc.live) return;
  const long long nz = a.nz, nr = a.nr;
  for (int k = 0; k < a.ncomp; k++) {
    double *T = a.jzero[k < 3 ? k : 0];
    if (!T) continue;
    const int c = a.comp[k];
    if (c == COMP_JZ) {
      if (pc.izA >= 0 && pc.izA <= nz && (pc.irh == nr - 1 || pc.irh == nr)) plus_zero(T, a.ldz, pc.izA, pc.irh - 1);
    } else if (c == COMP_JR) {
      if (pc.izh >= 0 && pc.izh <= nz && (pc.irA == 0 || pc.irA == nr - 1))
        plus_zero(T, a.ldz, pc.izh, pc.irA == 0 ? 0 : nr - 2);
    } else if (c == COMP_JPHI) {
      if (pc.izh >= 0 && pc.izh <= nz && (pc.irh == nr - 1 || pc.irh == nr)) plus_zero(T, a.ldz, pc.izh, pc.irh - 1);
    }
  }
}

// the targets' reset (src/integrate.c:169 zeroes J before the deposition),
// done by the key kernels to save a launch; the folds that add into them come later
__device__ inline void clear_targets(const KeyArgs &a, int64_t ip, int64_t stride) {
  for (int64_t q = ip; q < a.clear_n; q += stride)
    for (int c = 0; c < 3; c++)
      if (a.clear[c]) a.clear[c][q] = 0.0;
}

// folds a workgroup's smallest and largest key into one of the range slots
template <int NWAVES>
__device__ inline void key_range_fold(unsigned long long *krange, uint32_t kmin, uint32_t kmax) {
  for (int off = 32; off > 0; off >>= 1) {
    kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, off, 64));
    kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, off, 64));
  }
  __shared__ uint32_t wmin[NWAVES], wmax[NWAVES];
  if ((threadIdx.x & 63) == 0) {
    wmin[threadIdx.x >> 6] = kmin;
    wmax[threadIdx.x >> 6] = kmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NWAVES; w++) {
      kmin = min(kmin, wmin[w]);
      kmax = max(kmax, wmax[w]);
    }
    if (kmin <= kmax) {
      const int slot = blockIdx.x % kRangeSlots;
      atomicMin(krange + slot, (unsigned long long)kmin);
      atomicMax(krange + kRangeSlots + slot, (unsigned long long)kmax);
    }
  }
}

__global__ __launch_bounds__(256) void key_kernel(KeyArgs a) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip == 0) a.counters[9] = 0;
  clear_targets(a, ip, (int64_t)gridDim.x * blockDim.x);
  // smallest and largest key this lane writes (every lane takes part in the
  // wavefront's reduction below, so no lane returns early)
  uint32_t kmin = 0xffffffffu, kmax = 0u;
  const bool in = ip < a.n_ub;
  int64_t np = in ? *a.np : 0;
  const PartVals pv = load_part(a, ip, np);
  const PartCells pc = part_cells(a, pv);
  if (in && a.jzero[0]) dropped_j_zeros(a, pc);
  for (int k = 0; in && k < a.ncomp; k++) {
    double v;
    const uint32_t key = part_contrib(a, pc, pv, k, v);
    int64_t slot = (int64_t)k * a.n_ub + ip;
    a.keys[slot] = key;
    a.vals[slot] = key != a.discard ? v : 0.0;
    if (key != a.discard) {
      kmin = min(kmin, key);
      kmax = max(kmax, key);
    }
  }
  if (!a.krange) return;   // the sort covers the whole key space
  key_range_fold<4>(a.krange, kmin, kmax);
}

// the key kernel of area-weighted deposition: four slots per particle and
// component, particle-major (slot 4*ip + t of component k's block)
__global__ __launch_bounds__(256) void key_smear_kernel(KeyArgs a) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip == 0) a.counters[9] = 0;
  clear_targets(a, ip, (int64_t)gridDim.x * blockDim.x);
  uint32_t kmin = 0xffffffffu, kmax = 0u;
  const bool in = ip < a.n_ub;
  int64_t np = in ? *a.np : 0;
  const PartVals pv = load_part(a, ip, np);
  for (int k = 0; in && k < a.ncomp; k++) {
    const int c = a.comp[k];
    Smear4 sm;
    if (c == COMP_Q)
      smear_charge(a, pv, sm);
    else
      smear_current(a, pv, c, sm);
    const int64_t base = (int64_t)k * 4 * a.n_ub + 4 * ip;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      uint32_t key = a.discard;
      double v = 0.0;
      if (sm.iz[t] >= 0) {
        key = ((uint32_t)(sm.ir[t] * (a.nz + 1) + sm.iz[t]) << 2) | (uint32_t)k;
        v = c == COMP_Q ? -sm.v[t] : sm.v[t];
        kmin = min(kmin, key);
        kmax = max(kmax, key);
      }
      a.keys[base + t] = key;
      a.vals[base + t] = v;
    }
  }
  if (!a.krange) return;
  key_range_fold<4>(a.krange, kmin, kmax);
}

// ---- deposition in a fixed order of our own (runs with cell-sorted particles)
//
// Each wavefront covers 64 * PPL consecutive particle slots, PPL per lane.
// For each key its particles contribute to, every lane adds its own
// contributions in slot order, the lanes' sums are added in one fixed tree
// over the lane positions (lanes without one add +0), and the wavefront
// emits one (key, partial) pair.  A
// workgroup's partials are written to its region in wavefront order, the
// regions are packed in workgroup order (count scan, then a copy), and the
// packed list goes through the same stable sort and fold as the per-particle
// pairs, with its length read on the device.  Each cell's sum is thus the
// fold, in storage order, of its wavefronts' tree sums: an order fixed
// by the storage order alone, so a run's result does not depend on
// scheduling, on the runs batched with it or on other contexts.  With
// particles sorted by cell, a wavefront's 64 particles share one or two
// cells, so there are far fewer partials than particles.
constexpr int RW = 4;     // wavefronts per reduce workgroup
constexpr int PPL = 4;    // particle slots per lane
constexpr int RG = 64 * RW * PPL;   // particle slots per reduce workgroup (reduce_groups: 1024)
static_assert(RG == 1024, "reduce_groups() counts workgroups of 1024 slots");

// x of lane (lane - shift within its row, or another row's lane for the
// broadcasts); lanes the control does not write take +0
template <int CTRL, int ROWS>
__device__ inline double dpp_take(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWS, 0xF, true);
  return __hiloint2double(hi, lo);
}

// The 64 lanes' x added in one fixed tree, the total in lane 63: within each
// row of 16 lanes, steps of 1, 2, 4 and 8 lanes leave the row's sum in its
// last lane; then row 1 adds row 0's and row 3 adds row 2's (row_bcast:15),
// and rows 2-3 add lane 31's (row_bcast:31).  Cross-lane moves are DPP, so
// no LDS is involved.
__device__ inline double wave_sum_to_last(double x) {
  x += dpp_take<0x111, 0xF>(x);
  x += dpp_take<0x112, 0xF>(x);
  x += dpp_take<0x114, 0xF>(x);
  x += dpp_take<0x118, 0xF>(x);
  x += dpp_take<0x142, 0xA>(x);
  x += dpp_take<0x143, 0xC>(x);
  return x;
}

// Workgroup b of nb of one run's reduce deposition; keys are emitted plus
// key_base (0 for a run alone; the run's offset in a batch's key space).
__device__ inline void key_reduce_body(const KeyArgs &a, int64_t b, int64_t nb, uint32_t key_base) {
  __shared__ uint32_t lk[RW][64 * PPL * 3];
  __shared__ double lv[RW][64 * PPL * 3];
  __shared__ int wcnt[RW + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t gbase = b * RG;   // first particle slot of the workgroup
  if (b == 0 && threadIdx.x == 0) a.counters[9] = 0;
  clear_targets(a, b * blockDim.x + threadIdx.x, nb * blockDim.x);
  uint32_t kmin = 0xffffffffu, kmax = 0u;
  const int64_t np = gbase < a.n_ub ? *a.np : 0;
  // lane l holds slots base + j*64 + l of the wavefront's 64*PPL consecutive slots
  const int64_t base = gbase + (int64_t)w * 64 * PPL;
  uint32_t key[3][PPL];
  double val[3][PPL];
  PartVals pv[PPL];
#pragma unroll
  for (int j = 0; j < PPL; j++) pv[j] = load_part(a, base + j * 64 + lane, np);
#pragma unroll
  for (int j = 0; j < PPL; j++) {
    const int64_t ip = base + j * 64 + lane;
    const bool in = ip < a.n_ub;
    const PartCells pc = part_cells(a, pv[j]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      key[k][j] = a.discard;
      val[k][j] = 0.0;
      if (in && k < a.ncomp) key[k][j] = part_contrib(a, pc, pv[j], k, val[k][j]);
      if (key[k][j] != a.discard) {
        kmin = min(kmin, key[k][j]);
        kmax = max(kmax, key[k][j]);
      }
    }
    if (a.pvals && in) a.pvals[ip] = key[0][j] != a.discard ? val[0][j] : 0.0;   // for the charge sums
  }
  int cnt = 0;   // partials of this wavefront (uniform)
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (k >= a.ncomp) break;
    // each round takes the first pending key of the lowest lane that has one;
    // every lane adds its values with that key in slot order, and the lanes'
    // sums are added in the fixed tree
    auto any_left = [&]() {
      bool m = false;
#pragma unroll
      for (int j = 0; j < PPL; j++) m |= key[k][j] != a.discard;
      return m;
    };
    unsigned long long pending = __ballot(any_left());
    while (pending) {
      const int leader = __ffsll((long long)pending) - 1;
      uint32_t first = a.discard;
#pragma unroll
      for (int j = PPL - 1; j >= 0; j--)
        if (key[k][j] != a.discard) first = key[k][j];
      const uint32_t K = (uint32_t)__builtin_amdgcn_readlane((int)first, leader);
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < PPL; j++)
        if (key[k][j] == K) {
          s += val[k][j];
          key[k][j] = a.discard;
        }
      const double x = wave_sum_to_last(s);
      if (lane == 63) {
        lk[w][cnt] = K + key_base;
        lv[w][cnt] = x;
      }
      cnt++;
      pending = __ballot(any_left());
    }
  }
  if (lane == 0) wcnt[w] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < RW; k++) {
      const int c = wcnt[k];
      wcnt[k] = run;
      run += c;
    }
    wcnt[RW] = run;
    // workgroups past the particles only clear the targets
    if (gbase < a.n_ub) a.wg_count[b] = run;
  }
  __syncthreads();
  // the workgroup's region holds at most ncomp partials per particle slot
  const int64_t region = gbase * a.ncomp + wcnt[w];
  for (int i = lane; i < cnt; i += 64) {
    a.keys[region + i] = lk[w][i];
    a.vals[region + i] = lv[w][i];
  }
  if (a.krange && kmin <= kmax) {
    kmin += key_base;
    kmax += key_base;
  }
  if (a.krange) key_range_fold<RW>(a.krange, kmin, kmax);
}

__global__ __launch_bounds__(64 * RW) void key_reduce_kernel(KeyArgs a) { key_reduce_body(a, blockIdx.x, gridDim.x, 0u); }

// exclusive scan of the reduce workgroups' partial counts (one workgroup);
// the total goes to *total
__device__ inline void partial_scan_body(const int32_t *cnt, int64_t n, int64_t *off, int64_t *total) {
  // 1024 threads; chunks of 8192 counts, thread t taking counts 8t .. 8t+7 of
  // a chunk (their eight loads issued together), one block scan per chunk
  constexpr int V = 8;
  __shared__ int64_t ws[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < n; c0 += 1024 * V) {
    const int64_t t0 = c0 + (int64_t)threadIdx.x * V;
    int32_t v[V];
#pragma unroll
    for (int k = 0; k < V; k++) v[k] = t0 + k < n ? cnt[t0 + k] : 0;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < V; k++) s += v[k];
    int64_t incl = s;
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    int64_t before = carry, chunk = 0;
    for (int k = 0; k < 16; k++) {
      if (k < w) before += ws[k];
      chunk += ws[k];
    }
    __syncthreads();   // ws is rewritten by the next chunk
    int64_t run = before + incl - s;
#pragma unroll
    for (int k = 0; k < V; k++) {
      if (t0 + k < n) off[t0 + k] = run;
      run += v[k];
    }
    carry += chunk;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void partial_scan_kernel(const int32_t *cnt, int64_t n, int64_t *off,
                                                            int64_t *total) {
  partial_scan_body(cnt, n, off, total);
}

// packs each reduce workgroup's partials (region b of stride rstride) to off[b]
__global__ __launch_bounds__(256) void partial_pack_kernel(const uint32_t *kin, const double *vin, const int32_t *cnt,
                                                           const int64_t *off, int64_t rstride, uint32_t *kout,
                                                           double *vout) {
  const int64_t b = blockIdx.x;
  const int c = cnt[b];
  const int64_t src = b * rstride, dst = off[b];
  for (int i = threadIdx.x; i < c; i += 256) {
    kout[dst + i] = kin[src + i];
    vout[dst + i] = vin[src + i];
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
  const int64_t *n_dev; // the number of pairs on the device (null: n)
  int64_t *counters;    // [9]: number of long segments
  int32_t *long_list;   // first positions of segments longer than LONG_SEG
  double *target[4];
  const struct FoldTarget *tt;   // several runs' keys in one key space (null: one run, target[])
  int ntt;
};

// a run's place in a batch's key space: keys [key_base, next run's key_base)
struct FoldTarget {
  uint32_t key_base;
  int nz, ldz;
  double *target[4];
};


__device__ inline double *target_cell(const FoldArgs &a, uint32_t key) {
  if (a.tt) {
    int lo = 0, hi = a.ntt;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.tt[mid].key_base <= key)
        lo = mid;
      else
        hi = mid;
    }
    const FoldTarget &t = a.tt[lo];
    const uint32_t kl = key - t.key_base;
    const uint32_t cell = kl >> 2;
    const int iz = (int)(cell % (uint32_t)(t.nz + 1)), ir = (int)(cell / (uint32_t)(t.nz + 1));
    return t.target[kl & 3u] + fix(t.ldz, iz, ir);
  }
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

// pairs to fold: the host count, or the device count when there is one
__device__ inline int64_t fold_pairs(const FoldArgs &a) {
  if (!a.n_dev) return a.n;
  const int64_t m = *a.n_dev;
  return m < a.n ? (m > 0 ? m : 0) : a.n;
}

// the sorted positions [base, base + 64) of n, one per lane
__device__ inline void fold_window(const FoldArgs &a, const uint32_t *__restrict__ sk, const double *__restrict__ sv,
                                   int64_t n, int64_t base) {
  const int lane = threadIdx.x & 63;
  const int64_t p = base + lane;
  const bool valid = p < n;
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
  if (sL < n && kL != a.discard) {
    if (sL + LONG_SEG < n && sk[sL + LONG_SEG] == kL) {
      longL = true;
    } else {
      // sorted keys: the run ends within LONG_SEG positions of sL
      for (int64_t q = base + 64;; q += 64) {
        const int64_t i = q + lane;
        const unsigned long long dm = __ballot(i >= n || sk[i] != kL);
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

__global__ __launch_bounds__(256) void fold_kernel(FoldArgs a) {
  // the buffer the ranged sort left its result in
  const bool in1 = a.range && ranged_sort_in_k1(a.range);
  const uint32_t *__restrict__ sk = in1 ? a.sk1 : a.sk;
  const double *__restrict__ sv = in1 ? a.sv1 : a.sv;
  const int64_t n = fold_pairs(a), nwin = (n + 63) / 64, wpb = blockDim.x / 64;
  // one 64-position window per wavefront (with a device count, a capped grid
  // loops over the windows); the loop condition is uniform over the wavefront
  for (int64_t win = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6); win < nwin; win += (int64_t)gridDim.x * wpb)
    fold_window(a, sk, sv, n, win * 64);
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
  const int64_t n = fold_pairs(a);
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
      kk[u] = i < n ? sk[i] : ~key;
      v[u] = i < n ? sv[i] : 0.0;
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
          kk[u] = i < n ? sk[i] : ~key;
          v[u] = i < n ? sv[i] : 0.0;
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

// ---- re-sort of a cell-sort run's particles
//
// Key: the particle's half-cell, row-major in r (2*(r/dr) and 2*(z-zmin)/dz
// truncated, clamped to the grid); slots past np get the discard key and sort
// last.  The stable ranged radix sort orders (key, slot) pairs, and
// permute_kernel gathers every particle column into the second set in that
// order.  Any deterministic key would do: it only decides locality.
struct CellSortArgs {
  const double *z, *r;
  const int64_t *np;
  int64_t n_ub;
  double zmin, dz, dr;
  int nz, nr;
  uint32_t *keys;
  int64_t *idx;
  uint32_t discard;
  unsigned long long *krange;
};

__global__ __launch_bounds__(256) void cell_key_kernel(CellSortArgs a) {
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = ip < a.n_ub;
  const int64_t np = in ? *a.np : 0;
  uint32_t key = a.discard, kmin = 0xffffffffu, kmax = 0u;
  if (in && ip < np) {
    const double hz = 2.0 * ((a.z[ip] - a.zmin) / a.dz), hr = 2.0 * (a.r[ip] / a.dr);
    const long long zmax = 2LL * a.nz + 1, rmax = 2LL * a.nr + 1;
    // NaN fails every comparison and lands in half-cell 0
    const long long iz = hz >= 0 ? (hz < (double)zmax ? (long long)hz : zmax) : 0;
    const long long ir = hr >= 0 ? (hr < (double)rmax ? (long long)hr : rmax) : 0;
    key = (uint32_t)(ir * (zmax + 1) + iz);
    kmin = kmax = key;
  }
  if (in) {
    a.keys[ip] = key;
    a.idx[ip] = ip;
  }
  key_range_fold<4>(a.krange, kmin, kmax);
}

struct PermArgs {
  const double *src[16];
  double *dst[16];
  const int32_t *pid;
  int32_t *pid2;
  const int16_t *gen;
  int16_t *gen2;
  const uint32_t *st;
  uint32_t *st2;
  const int64_t *np;
  int64_t n_ub;
  const int64_t *idx0, *idx1;   // the sort's two value buffers (slot numbers)
  const uint64_t *range;
};

__global__ __launch_bounds__(256) void permute_kernel(PermArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_ub || i >= *a.np) return;
  const int64_t *ix = ranged_sort_in_k1(a.range) ? a.idx1 : a.idx0;
  const int64_t j = ix[i];
#pragma unroll
  for (int k = 0; k < 16; k++) a.dst[k][i] = a.src[k][j];
  a.pid2[i] = a.pid[j];
  a.gen2[i] = a.gen[j];
  a.st2[i] = a.st[j];
}

struct PtrTable {
  double *p[16];
};

__global__ void set_ptrs_kernel(PtrTable t, double **dst) {
  if (threadIdx.x < 16) dst[threadIdx.x] = t.p[threadIdx.x];
}

// ---- the reduce deposition of several cell-sort runs in one set of launches
//
// Each run's keys are offset by the run's base in a batch key space (the
// sum of the runs' key spaces), its workgroups' partials are packed run by
// run into one list, and one ranged sort and one fold over that list find
// each key's run by its base.  A run's partials, their order within each key
// and the fold's additions are those of the run's own reduce deposition, so
// every run gets the bits it gets alone.
struct ReduceRun {
  KeyArgs k;
  uint32_t key_base;
  int64_t nb;          // workgroups of the run's reduce launch share (particles, or the clear when wider)
  int64_t nwg;         // workgroups holding particle slots
  int64_t rstride;     // region stride of the run's partials (RG * ncomp)
  int64_t *wg_off;
  int64_t *total;      // the run's partial count
};

__device__ inline int ragged_run_d(const int32_t *first, int n, int b) {
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

__global__ __launch_bounds__(64 * RW) void key_reduce_multi_kernel(const ReduceRun *__restrict__ T,
                                                                  const int32_t *__restrict__ first, int nruns) {
  const int r = ragged_run_d(first, nruns, (int)blockIdx.x);
  key_reduce_body(T[r].k, (int64_t)blockIdx.x - first[r], T[r].nb, T[r].key_base);
}

__global__ __launch_bounds__(1024) void partial_scan_multi_kernel(const ReduceRun *__restrict__ T) {
  const ReduceRun &R = T[blockIdx.x];
  partial_scan_body(R.k.wg_count, R.nwg, R.wg_off, R.total);
}

// run_off[r]: the first packed position of run r; *m: all partials; counters[9] cleared for the fold
__global__ void run_offsets_kernel(const ReduceRun *__restrict__ T, int nruns, int64_t *run_off, int64_t *m,
                                   int64_t *counters) {
  if (threadIdx.x != 0) return;
  int64_t run = 0;
  for (int r = 0; r < nruns; r++) {
    run_off[r] = run;
    run += *T[r].total;
  }
  *m = run;
  counters[9] = 0;
}

__global__ __launch_bounds__(256) void partial_pack_multi_kernel(const ReduceRun *__restrict__ T,
                                                                 const int32_t *__restrict__ first, int nruns,
                                                                 const int64_t *run_off, uint32_t *kout, double *vout) {
  const int r = ragged_run_d(first, nruns, (int)blockIdx.x);
  const int64_t b = (int64_t)blockIdx.x - first[r];
  const ReduceRun &R = T[r];
  const int c = R.k.wg_count[b];
  const int64_t src = b * R.rstride, dst = run_off[r] + R.wg_off[b];
  for (int i = threadIdx.x; i < c; i += 256) {
    kout[dst + i] = R.k.keys[src + i];
    vout[dst + i] = R.k.vals[src + i];
  }
}

}  // namespace

int deposit_slots_per_particle(uint32_t modes) { return (modes & (SPF_RADIAL_SMEAR | SPF_LONGIT_SMEAR)) ? 4 : 1; }

int deposit_cell_bits(int nz, int nr) {
  uint64_t cells = (uint64_t)(nz + 1) * (uint64_t)(nr + 1);
  int b = 1;
  while ((1ull << b) < cells) b++;
  return b;
}

// the key kernels' arguments for the listed components of run R
static KeyArgs make_key_args(Run &R, const int *comps, int ncomp, double *const *targets, bool clear_first,
                             bool zero_terms, uint32_t discard, int cell_bits) {
  const int64_t n_ub = R.np_ub;
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
  k.pvals = nullptr;
  k.wg_count = nullptr;
  for (int i = 0; i < 3; i++) k.need_v[i] = 0;
  for (int i = 0; i < ncomp; i++)
    if (comps[i] <= COMP_JPHI) k.need_v[comps[i]] = 1;
  k.smear_r = (R.d.modes & SPF_RADIAL_SMEAR) != 0;
  k.smear_z = (R.d.modes & SPF_LONGIT_SMEAR) != 0;
  k.tpp = deposit_slots_per_particle(R.d.modes);
  for (int i = 0; i < 3; i++) k.jzero[i] = nullptr;
  if (!clear_first && ncomp > 0 && comps[0] != COMP_Q)
    for (int i = 0; i < ncomp && i < 3; i++) k.jzero[i] = targets[i];
  // smeared charge removal keeps its zero-weight terms in the ordered fold
  if (k.tpp != 1) {
    k.zero_terms = 0;
    k.zero_target = nullptr;
  }
  return k;
}

// The reduce deposition of cell-sort runs: per-wavefront tree-sum partials,
// packed in workgroup order, then the ranged sort and the fold over the
// packed list with its length read on the device (counters[10]).
static hipError_t reduce_scatter(Run &R, KeyArgs k, int ncomp, double *const *targets, hipStream_t s,
                                 bool clear_first, double *sums, int key_bits, uint32_t discard) {
  const int64_t n_ub = R.np_ub, n = n_ub * ncomp, nwg = reduce_groups(n_ub);
  k.keys = R.keys;
  k.vals = R.cz;
  k.pvals = sums ? R.scz : nullptr;
  k.wg_count = R.wg_count;
  k.krange = reinterpret_cast<unsigned long long *>(R.krange);
  int64_t kblocks = nwg;
  if (clear_first) kblocks = std::max<int64_t>(kblocks, std::min<int64_t>(8192, (R.cells + 64 * RW - 1) / (64 * RW)));
  hipLaunchKernelGGL(key_range_reset_kernel, dim3(1), dim3(kRangeSlots), 0, s, k.krange);
  hipLaunchKernelGGL(key_reduce_kernel, dim3((unsigned)kblocks), dim3(64 * RW), 0, s, k);
  hipLaunchKernelGGL(key_range_final_kernel, dim3(1), dim3(kRangeSlots), 0, s, k.krange);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (sums && (e = launch_sum2_vec(R.scz, n_ub, R.red, R.red_n, sums, s)) != hipSuccess) return e;
  int64_t *total = R.counters + 10;
  hipLaunchKernelGGL(partial_scan_kernel, dim3(1), dim3(1024), 0, s, R.wg_count, nwg, R.wg_off, total);
  hipLaunchKernelGGL(partial_pack_kernel, dim3((unsigned)nwg), dim3(256), 0, s, R.keys, R.cz, R.wg_count, R.wg_off,
                     (int64_t)RG * ncomp, R.skeys, R.scz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const uint64_t *range = reinterpret_cast<const uint64_t *>(R.krange + 2 * kRangeSlots);
  e = radix_sort_pairs_ranged(R.skeys, R.scz, R.keys, R.cz, n, key_bits, range, discard, R.sort_tmp,
                              R.sort_tmp_bytes, s, total);
  if (e != hipSuccess) return e;
  FoldArgs f;
  f.nz = R.d.nz;
  f.ldz = R.ldz;
  f.cell_bits = key_bits - 2;
  f.n = n;
  f.discard = discard;
  f.sk = R.skeys;
  f.sv = R.scz;
  f.sk1 = R.keys;
  f.sv1 = R.cz;
  f.range = range;
  f.n_dev = total;
  f.counters = R.counters;
  f.long_list = R.long_list;
  f.tt = nullptr;
  f.ntt = 0;
  for (int i = 0; i < 4; i++) f.target[i] = i < ncomp ? targets[i] : nullptr;
  hipLaunchKernelGGL(fold_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, f);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(fold_long_kernel, dim3((unsigned)std::min<int64_t>(n / LONG_SEG + 1, 2048)), dim3(64), 0, s, f);
  return hipGetLastError();
}

// ordered scatter-add of the listed components into their target arrays; with
// clear_first (at most three targets) the targets are set to +0 first
static hipError_t ordered_scatter(Run &R, const int *comps, int ncomp, double *const *targets, hipStream_t s,
                                  bool clear_first = false, bool zero_terms = false, double *sums = nullptr) {
  const int64_t n_ub = R.np_ub;
  if (clear_first && (n_ub == 0 || ncomp == 0))
    return launch_fill3(ncomp > 0 ? targets[0] : nullptr, ncomp > 1 ? targets[1] : nullptr,
                        ncomp > 2 ? targets[2] : nullptr, R.cells, 0.0, s);
  if (n_ub == 0 || ncomp == 0) return hipSuccess;
  const int tpp = deposit_slots_per_particle(R.d.modes);
  const int64_t n = n_ub * ncomp * tpp;
  const int cell_bits = deposit_cell_bits(R.d.nz, R.d.nr);
  const int key_bits = cell_bits + 2;
  const uint32_t discard = (1u << key_bits) - 1u;   // all-ones cell with slot 3: never a real key
  KeyArgs k = make_key_args(R, comps, ncomp, targets, clear_first, zero_terms, discard, cell_bits);
  // cell-sort runs use the reduce deposition, except when smeared: a particle
  // then has several targets per component, and the ordered path serves
  if (R.cell_sort > 0 && tpp == 1)
    return reduce_scatter(R, k, ncomp, targets, s, clear_first, sums, key_bits, discard);
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
  if (tpp == 4)
    hipLaunchKernelGGL(key_smear_kernel, dim3((unsigned)kblocks), dim3(256), 0, s, k);
  else
    hipLaunchKernelGGL(key_kernel, dim3((unsigned)kblocks), dim3(256), 0, s, k);
  if (ranged) hipLaunchKernelGGL(key_range_final_kernel, dim3(1), dim3(kRangeSlots), 0, s, k.krange);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the first component's values in particle order (tpp per particle), before
  // the sort reuses their buffer
  if (sums && (e = launch_sum2_vec(R.cz, n_ub * tpp, R.red, R.red_n, sums, s)) != hipSuccess) return e;
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
  f.n_dev = nullptr;
  f.counters = R.counters;
  f.long_list = R.long_list;
  f.tt = nullptr;
  f.ntt = 0;
  for (int i = 0; i < 4; i++) f.target[i] = i < ncomp ? targets[i] : nullptr;
  hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // at most n / LONG_SEG long segments
  int64_t nlong = n / LONG_SEG + 1;
  hipLaunchKernelGGL(fold_long_kernel, dim3((unsigned)std::min<int64_t>(nlong, 8192)), dim3(64), 0, s, f);
  return hipGetLastError();
}


// The J reset and deposition of several cell-sort runs (space charge, at
// least one component, particles present) in one set of launches.
hipError_t deposit_currents_multi(spf_ctx *c, Run *const *runs, int n) {
  hipStream_t s = c->stream;
  std::vector<ReduceRun> rr(n);
  std::vector<FoldTarget> ft(n);
  std::vector<int32_t> first(n + 1), first2(n + 1);
  uint64_t key_base = 0;
  int64_t total_pairs = 0, blocks = 0, pblocks = 0;
  for (int i = 0; i < n; i++) {
    Run &R = *runs[i];
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
    const int cell_bits = deposit_cell_bits(R.d.nz, R.d.nr);
    ReduceRun &q = rr[i];
    q.k = make_key_args(R, comps, nc, targets, true, false, 0xffffffffu, cell_bits);
    q.k.wg_count = R.wg_count;
    q.key_base = (uint32_t)key_base;
    q.nwg = reduce_groups(R.np_ub);
    q.nb = std::max<int64_t>(q.nwg, std::min<int64_t>(2048, (R.cells + 64 * RW - 1) / (64 * RW)));
    q.rstride = (int64_t)RG * nc;
    q.wg_off = R.wg_off;
    q.total = R.counters + 10;
    ft[i].key_base = (uint32_t)key_base;
    ft[i].nz = R.d.nz;
    ft[i].ldz = R.ldz;
    for (int k = 0; k < 4; k++) ft[i].target[k] = k < nc ? targets[k] : nullptr;
    first[i] = (int32_t)blocks;
    first2[i] = (int32_t)pblocks;
    blocks += q.nb;
    pblocks += q.nwg;
    total_pairs += R.np_ub * nc;
    key_base += 4ull * (uint64_t)(R.d.nz + 1) * (uint64_t)(R.d.nr + 1);
  }
  first[n] = (int32_t)blocks;
  first2[n] = (int32_t)pblocks;
  if (key_base >= 0xffffffffull || blocks >= (1ll << 31) || total_pairs >= (1ll << 31)) return hipErrorInvalidValue;
  if (deposit_scratch(c, total_pairs, n) != SPF_OK) return hipErrorOutOfMemory;
  unsigned long long *krange = reinterpret_cast<unsigned long long *>(c->dep_krange);
  for (int i = 0; i < n; i++) rr[i].k.krange = krange;
  // one table: the runs, both block maps, the fold targets
  const size_t o1 = (sizeof(ReduceRun) * (size_t)n + 15) / 16 * 16;
  const size_t o2 = o1 + (sizeof(int32_t) * (size_t)(n + 1) + 15) / 16 * 16;
  const size_t o3 = o2 + (sizeof(int32_t) * (size_t)(n + 1) + 15) / 16 * 16;
  const size_t bytes = o3 + sizeof(FoldTarget) * (size_t)n;
  void *host, *dev;
  if (batch_table(c, bytes, &host, &dev) != SPF_OK) return hipErrorOutOfMemory;
  memcpy(host, rr.data(), sizeof(ReduceRun) * (size_t)n);
  memcpy((char *)host + o1, first.data(), sizeof(int32_t) * (size_t)(n + 1));
  memcpy((char *)host + o2, first2.data(), sizeof(int32_t) * (size_t)(n + 1));
  memcpy((char *)host + o3, ft.data(), sizeof(FoldTarget) * (size_t)n);
  hipError_t e = batch_table_upload(c, host, dev, bytes);
  if (e != hipSuccess) return e;
  const ReduceRun *T = (const ReduceRun *)dev;
  const int32_t *F1 = (const int32_t *)((const char *)dev + o1), *F2 = (const int32_t *)((const char *)dev + o2);
  const FoldTarget *TT = (const FoldTarget *)((const char *)dev + o3);
  int64_t *misc = c->dep_misc;   // [0] packed partials, [9] long segments
  hipLaunchKernelGGL(key_range_reset_kernel, dim3(1), dim3(kRangeSlots), 0, s, krange);
  hipLaunchKernelGGL(key_reduce_multi_kernel, dim3((unsigned)blocks), dim3(64 * RW), 0, s, T, F1, n);
  hipLaunchKernelGGL(key_range_final_kernel, dim3(1), dim3(kRangeSlots), 0, s, krange);
  hipLaunchKernelGGL(partial_scan_multi_kernel, dim3((unsigned)n), dim3(1024), 0, s, T);
  hipLaunchKernelGGL(run_offsets_kernel, dim3(1), dim3(64), 0, s, T, n, c->dep_run_off, misc, misc);
  if (pblocks > 0)
    hipLaunchKernelGGL(partial_pack_multi_kernel, dim3((unsigned)pblocks), dim3(256), 0, s, T, F2, n, c->dep_run_off,
                       c->dep_k0, c->dep_v0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int key_bits = 1;
  while (key_bits < 32 && (1ull << key_bits) <= key_base) key_bits++;
  const uint64_t *range = reinterpret_cast<const uint64_t *>(c->dep_krange + 2 * kRangeSlots);
  e = radix_sort_pairs_ranged(c->dep_k0, c->dep_v0, c->dep_k1, c->dep_v1, total_pairs, key_bits, range, 0xffffffffu,
                              c->dep_tmp, c->dep_tmp_bytes, s, misc);
  if (e != hipSuccess) return e;
  FoldArgs f;
  f.nz = 0;
  f.ldz = 0;
  f.cell_bits = 0;
  f.n = total_pairs;
  f.discard = 0xffffffffu;
  f.sk = c->dep_k0;
  f.sv = c->dep_v0;
  f.sk1 = c->dep_k1;
  f.sv1 = c->dep_v1;
  f.range = range;
  f.n_dev = misc;
  f.counters = misc;
  f.long_list = c->dep_long;
  for (int i = 0; i < 4; i++) f.target[i] = nullptr;
  f.tt = TT;
  f.ntt = n;
  hipLaunchKernelGGL(fold_kernel, dim3((unsigned)std::min<int64_t>((total_pairs + 255) / 256, 4096)), dim3(256), 0, s,
                     f);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(fold_long_kernel, dim3((unsigned)std::min<int64_t>(total_pairs / LONG_SEG + 1, 2048)), dim3(64),
                     0, s, f);
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

// Qreal of src/poisson_correction.c:220-239 summed in particle order from +0,
// one rounded addition at a time.  Without smearing fb = fl = 1 and the other
// three terms weigh zero: they add +-0, which never changes a sum that
// starts at +0.  With smearing every unblocked term is added, in the order
// (iz, ir), (iz+1, ir), (iz, ir+1), (iz+1, ir+1).
__global__ __launch_bounds__(64) void qreal_serial_kernel(KeyArgs a, double *out) {
  __shared__ double buf[64 * 4];
  __shared__ unsigned char tk[64 * 4];
  const int lane = threadIdx.x;
  const int64_t np = *a.np;
  const bool smear = a.smear_r || a.smear_z;
  double s = 0.0;
  for (int64_t c = 0; c < np; c += 64) {
    const int64_t ip = c + lane;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    bool take[4] = {false, false, false, false};
    if (smear) {
      PartVals pv{false, 0, 0, 0, 0, 0, 0};
      if (ip < np) {
        pv.live = (a.status[ip] & SPF_PART_ACTIVE) != 0;
        pv.Q = a.Q[ip];
        pv.z = a.z[ip];
        pv.r = a.r[ip];
      }
      Smear4 sm;
      smear_charge(a, pv, sm);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        take[k] = sm.iz[k] >= 0;
        t[k] = sm.v[k];
      }
    } else if (ip < np && (a.status[ip] & SPF_PART_ACTIVE)) {
      const double th = (a.z[ip] - a.zmin) / a.dz + 0.5, yh = a.r[ip] / a.dr + 0.5;
      if (th > -1.0 && th < a.nz + 1.0 && yh > -1.0 && yh < a.nr + 1.0) {
        const long long iz = (long long)th, ir = (long long)yh;
        if (iz >= 0 && iz < a.nz && ir >= 0 && ir < a.nr) {
          const int16_t f = a.flags[fix(a.ldz, (int)iz, (int)ir)];
          take[0] = !((f & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (iz == 0 && a.left_neu) ||
                      (iz >= a.nz - 1 && a.right_neu) || (ir >= a.nr - 1 && a.upper_neu));
          t[0] = a.Q[ip] * 1.0 * 1.0;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      buf[4 * lane + k] = t[k];
      tk[4 * lane + k] = take[k];
    }
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 64 * 4; k++)
        if (tk[k]) s += buf[k];
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
  k.smear_r = (R.d.modes & SPF_RADIAL_SMEAR) != 0;
  k.smear_z = (R.d.modes & SPF_LONGIT_SMEAR) != 0;
  k.tpp = deposit_slots_per_particle(R.d.modes);
  hipLaunchKernelGGL(qreal_serial_kernel, dim3(1), dim3(64), 0, s, k, out);
  return hipGetLastError();
}

hipError_t sort_particles(Run &R, hipStream_t s) {
  const int64_t n_ub = R.np_ub;
  if (R.cell_sort <= 0 || n_ub == 0) return hipSuccess;
  CellSortArgs a;
  a.z = R.P[1];
  a.r = R.P[2];
  a.np = R.counters;
  a.n_ub = n_ub;
  a.zmin = R.d.zmin;
  a.dz = R.d.dz;
  a.dr = R.d.dr;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.keys = R.keys;
  a.idx = reinterpret_cast<int64_t *>(R.cz);
  const uint64_t nkeys = (uint64_t)(2 * R.d.nz + 2) * (uint64_t)(2 * R.d.nr + 2);
  int key_bits = 1;
  while ((1ull << key_bits) <= nkeys) key_bits++;
  a.discard = 0xffffffffu;
  a.krange = reinterpret_cast<unsigned long long *>(R.krange);
  hipLaunchKernelGGL(key_range_reset_kernel, dim3(1), dim3(kRangeSlots), 0, s, a.krange);
  hipLaunchKernelGGL(cell_key_kernel, dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(key_range_final_kernel, dim3(1), dim3(kRangeSlots), 0, s, a.krange);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t *range = reinterpret_cast<const uint64_t *>(R.krange + 2 * kRangeSlots);
  e = radix_sort_pairs_ranged(R.keys, R.cz, R.skeys, R.scz, n_ub, key_bits, range, a.discard, R.sort_tmp,
                              R.sort_tmp_bytes, s);
  if (e != hipSuccess) return e;
  PermArgs p;
  for (int k = 0; k < 16; k++) {
    p.src[k] = R.P[k];
    p.dst[k] = R.P2[k];
  }
  p.pid = R.pid;
  p.pid2 = R.pid2;
  p.gen = R.gen;
  p.gen2 = R.gen2;
  p.st = R.status;
  p.st2 = R.status2;
  p.np = R.counters;
  p.n_ub = n_ub;
  p.idx0 = reinterpret_cast<const int64_t *>(R.cz);
  p.idx1 = reinterpret_cast<const int64_t *>(R.scz);
  p.range = range;
  hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, s, p);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the gathered set becomes the current one; kernels take the pointers as
  // arguments, and the device copy used by emission is rewritten in order
  PtrTable t;
  for (int k = 0; k < 16; k++) {
    std::swap(R.P[k], R.P2[k]);
    t.p[k] = R.P[k];
  }
  std::swap(R.pid, R.pid2);
  std::swap(R.gen, R.gen2);
  std::swap(R.status, R.status2);
  hipLaunchKernelGGL(set_ptrs_kernel, dim3(1), dim3(64), 0, s, t, R.Pdev);
  return hipGetLastError();
}

}  // namespace spf
