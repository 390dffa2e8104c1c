// spf_push.hip -- relativistic particle push with staggered field gather,
// loss detection and screen crossing (src/advance_particles.c:67-575), the
// ordered per-step event rows, and particle compaction.
//
// One thread per particle, particles as structure-of-arrays in the
// reference's index order.  Events (losses, screen crossings) are flagged per
// particle; an ordered compaction turns them into an index-ordered list, so
// screen and lost rows come out in the same order as the reference's serial
// loops regardless of scheduling.
#include <climits>
#include <cstring>

#include "spf_internal.h"

namespace spf {

namespace {

constexpr uint32_t ACTIVE = SPF_PART_ACTIVE, LOST = SPF_PART_LOST;
constexpr uint32_t EV_LOST = 1u, EV_RNEG = 2u;  // screen s crossing: bit 2+s

// C conversion double -> long truncates toward zero; out-of-range values and
// NaN give LONG_MIN, as the x86-64 conversion instruction does.  Grid indices
// are far inside 32 bits, where the one-instruction 32-bit conversion
// truncates to the same value; the 64-bit sequence runs only for the rest.
__device__ inline long long trunc_l(double x) {
  if (__builtin_expect(__builtin_fabs(x) < 2147483648.0, 1)) return (long long)(int)x;
  if (!(x > -9.2233720368547758e18 && x < 9.2233720368547758e18)) return LLONG_MIN;
  return (long long)x;
}

__device__ inline double sqr(double x) { return x * x; }

// OPEN: the particle's cell is an open cell (PushArgs::open), where every flag
// bit the loss test and the gather read is clear, so fl() is the constant 0
// and those tests compile away; the arithmetic is the same code.
template <bool OPEN>
struct G {
  const PushArgs &a;
  __device__ int16_t fl(long long iz, long long ir) const {
    return OPEN ? (int16_t)0 : a.flags[fix(a.ldz, (int)iz, (int)ir)];
  }
  __device__ double at(const double *p, long long iz, long long ir) const { return p[fix(a.ldz, (int)iz, (int)ir)]; }
  // imposed E: while the run has never set one its arrays hold +0.0, so the
  // literal +0.0 gives the same sums without the loads
  __device__ double imp(const double *p, long long iz, long long ir) const { return a.imp_e ? at(p, iz, ir) : 0.0; }

  // src/advance_particles.c:31-65
  __device__ short loss_code(long long iz, long long ir) const {
    // an inactive particle at r < -rmax reaches here with ir > nr, where the
    // reference reads outside its flag array; such a point counts as metal
    if (iz < 0 || iz > a.nz || ir < 0 || ir > a.nr) return 3;
    short lossCode = 0;
    int16_t f = fl(iz, ir);
    if (f & SPF_FL_IS_INSIDE)
      lossCode = 3;
    else if (f & SPF_FL_IS_SURFACE) {
      long long iz1 = iz + 1;
      if (iz1 > a.nz)
        lossCode = 4;
      else {
        long long ir1 = ir + 1;
        if (ir1 > a.nr)
          lossCode = 5;
        else if ((fl(iz1, ir1) & 3) && (fl(iz, ir1) & 3) && (fl(iz1, ir) & 3))
          lossCode = 6;
      }
    }
    return lossCode;
  }

  // src/advance_particles.c:622-890
  __device__ void interp(double z, double r, double &EzR, double &ErR, double &EphiR, double &BzR, double &BrR,
                         double &BphiR) const {
    long long iz, ir, iz1, ir1, nz = a.nz, nr = a.nr;
    double fr, fl_, fa, fb, dz = a.dz, dr = a.dr, zmin = a.zmin;
    double Ez = 0, Er = 0, Ephi = 0, Bz = 0, Br = 0, Bphi = 0;
    // -- Ez and Br
    if (a.longit) {
      fr = (z - zmin) / dz + 0.5;
      iz = trunc_l(fr);
      fr -= iz;
      fl_ = 1 - fr;
    } else {
      iz = trunc_l((z - zmin) / dz + 1);
      fl_ = 1;
      fr = 0;
    }
    if (a.radial) {
      fa = r / dr;
      ir = trunc_l(fa);
      fa -= ir;
      fb = 1 - fa;
    } else {
      ir = trunc_l(r / dr + 0.5);
      fa = 0;
      fb = 1;
    }
    if (iz >= 0 && iz < nz && ir >= 0 && ir < nr) {
      if ((iz1 = iz + 1) >= nz) iz1 = iz;
      if ((ir1 = ir + 1) >= nr) ir1 = ir;
      if (fl(iz, ir + 1) & SPF_FL_ER_ZERO && fl(iz, ir) & SPF_FL_ER_ZERO) {
        if (iz > 0 && (fl(iz - 1, ir) || fl(iz - 1, ir + 1)))
          iz = iz1;
        else
          iz1 = iz;
      } else if (fl(iz1, ir + 1) & SPF_FL_ER_ZERO && fl(iz1, ir) & SPF_FL_ER_ZERO) {
        iz1 = iz;
      }
      if (a.tm)
        Ez = fb * (fl_ * (at(a.Ez, iz, ir) + imp(a.EzImp, iz, ir)) + fr * (at(a.Ez, iz1, ir) + imp(a.EzImp, iz1, ir))) +
             fa * (fl_ * (at(a.Ez, iz, ir1) + imp(a.EzImp, iz, ir1)) + fr * (at(a.Ez, iz1, ir1) + imp(a.EzImp, iz1, ir1)));
      if (r != 0 && a.te)
        Br = fb * (fl_ * at(a.Br2, iz, ir) + fr * at(a.Br2, iz1, ir)) + fa * (fl_ * at(a.Br2, iz, ir1) + fr * at(a.Br2, iz1, ir1));
    }
    // -- Er and Bz
    if (a.longit) {
      fr = (z - zmin) / dz;
      iz = trunc_l(fr);
      fr -= iz;
      fl_ = 1 - fr;
    } else {
      iz = trunc_l((z - zmin) / dz + 0.5);
      fr = 0;
      fl_ = 1;
    }
    if (a.radial) {
      fa = r / dr + 0.5;
      ir = trunc_l(fa);
      fa -= ir;
      fb = 1 - fa;
    } else {
      ir = trunc_l(r / dr + 1);
      fa = 0;
      fb = 1;
    }
    if (iz >= 0 && iz < nz && ir >= 0) {
      if (ir >= nr) ir = nr - 1;
      iz1 = iz + 1;
      ir1 = ir + 1;
      if (fl(iz + 1, ir) & SPF_FL_EZ_ZERO && fl(iz, ir) & SPF_FL_EZ_ZERO)
        ir = ir1;
      else if (fl(iz + 1, ir + 1) & SPF_FL_EZ_ZERO && fl(iz, ir + 1) & SPF_FL_EZ_ZERO)
        ir1 = ir;
      if (r != 0 && a.tm)
        Er = fl_ * (fb * (at(a.Er, iz, ir) + imp(a.ErImp, iz, ir)) + fa * (at(a.Er, iz, ir1) + imp(a.ErImp, iz, ir1))) +
             fr * (fb * (at(a.Er, iz1, ir) + imp(a.ErImp, iz1, ir)) + fa * (at(a.Er, iz1, ir1) + imp(a.ErImp, iz1, ir1)));
      if (a.te)
        Bz = fl_ * (fb * at(a.Bz2, iz, ir) + fa * at(a.Bz2, iz, ir1)) + fr * (fb * at(a.Bz2, iz1, ir) + fa * at(a.Bz2, iz1, ir1));
    }
    // -- Ephi
    if (a.longit) {
      fr = (z - zmin) / dz;
      iz = trunc_l(fr);
      fr -= iz;
      fl_ = 1 - fr;
    } else {
      iz = trunc_l((z - zmin) / dz + 0.5);
      fr = 0;
      fl_ = 1;
    }
    if (a.radial) {
      fa = r / dr;
      ir = trunc_l(fa);
      fa -= ir;
      fb = 1 - fa;
    } else {
      ir = trunc_l(r / dr + 0.5);
      fa = 0;
      fb = 1;
    }
    if (r != 0 && a.te && iz >= 0 && iz < nz && ir >= 0 && ir < nr) {
      if ((iz1 = iz + 1) >= nz) iz1 = iz;
      if ((ir1 = ir + 1) >= nr) ir1 = ir;
      Ephi = fl_ * fb * at(a.Ephi, iz, ir) + fr * fb * at(a.Ephi, iz1, ir) + fl_ * fa * at(a.Ephi, iz, ir1) +
             fr * fa * at(a.Ephi, iz1, ir1);
    }
    // -- Bphi at time t
    if (a.tm) {
      if (a.longit) {
        fr = (z - zmin) / dz + 0.5;
        iz = trunc_l(fr);
        fr -= iz;
        fl_ = 1 - fr;
      } else {
        iz = trunc_l((z - zmin) / dz + 1);
        fr = 0;
        fl_ = 1;
      }
      if (a.radial) {
        fa = r / dr + 0.5;
        ir = trunc_l(fa);
        fa -= ir;
        fb = 1 - fa;
      } else {
        ir = trunc_l(r / dr + 1);
        fa = 0;
        fb = 1;
      }
      // ir < 0 only for NaN or astronomically large r, where the reference
      // would index outside its arrays; the device skips the component.
      if (r != 0 && iz >= 0 && iz < nz && ir >= 0) {
        if ((iz1 = iz + 1) >= nz) iz1 = iz;
        if (ir >= nr) ir = nr - 1;
        if ((ir1 = ir + 1) >= nr) ir1 = ir;
        Bphi = fl_ * fb * at(a.Bphi2, iz, ir) + fr * fb * at(a.Bphi2, iz1, ir) + fa * fl_ * at(a.Bphi2, iz, ir1) +
               fr * fa * at(a.Bphi2, iz1, ir1);
      }
    }
    // constants, then the on-axis expansion's +0.0 (none defined), :829-844
    Ez += a.cst[0];
    Er += a.cst[1];
    Ephi += a.cst[2];
    Bz += a.cst[3];
    Br += a.cst[4];
    Bphi += a.cst[5];
    Ez += 0.0;
    Er += 0.0;
    Bphi += 0.0;
    if (a.turnoff[0]) Ez = 0;
    if (a.turnoff[1]) Er = 0;
    if (a.turnoff[2]) Ephi = 0;
    if (a.turnoff[3]) Bz = 0;
    if (a.turnoff[4]) Br = 0;
    if (a.turnoff[5]) Bphi = 0;
    if (a.has_sol) {
      fr = (z - zmin) / dz;
      iz = trunc_l(fr);
      fa = r / dr;
      ir = trunc_l(fa);
      if (ir >= 0 && iz >= 0 && ir < nr && iz < nz) {
        fa -= ir;
        fb = 1 - fa;
        fr -= iz;
        fl_ = 1 - fr;
        Bz += fl_ * (fb * at(a.BzImp, iz, ir) + fa * at(a.BzImp, iz, ir + 1)) +
              fr * (fb * at(a.BzImp, iz + 1, ir) + fa * at(a.BzImp, iz + 1, ir + 1));
        Br += fl_ * (fb * at(a.BrImp, iz, ir) + fa * at(a.BrImp, iz, ir + 1)) +
              fr * (fb * at(a.BrImp, iz + 1, ir) + fa * at(a.BrImp, iz + 1, ir + 1));
      }
    }
    EzR = Ez;
    ErR = Er;
    EphiR = Ephi;
    BzR = Bz;
    BrR = Br;
    BphiR = Bphi;
  }
};

// src/advance_particles.c:131-339 for one particle; returns whether the
// particle has an event (loss or screen crossing) this step
__device__ inline bool push_one(const PushArgs &a, int64_t ip, int64_t np) {
  G<false> g{a};
  double dt = a.dt, dz = a.dz, dr = a.dr, zmin = a.zmin, zmax = a.zmax, rmax = a.rmax;
  double z = a.z[ip], r = a.r[ip];
  const double zold = z, rold = r;
  // The extents are updated with the old position here and with the new one
  // below when the particle stays active (:151-158, :253-260); both updates
  // are max/min comparisons, so they are applied together at the end, in
  // the same order, and the extents need no registers through the gather.
  double vz = a.vz[ip], vr = a.vr[ip];
  z += vz * dt;
  r += vr * dt;
  const uint32_t st0 = a.status[ip];
  uint32_t st = st0;
  double adjustFactor = 1;
  if (!(st & ACTIVE)) {
    long long iz = trunc_l((z - zmin) / dz);
    long long ir = trunc_l(fabs(r) / dr);
    bool inside = z >= a.zmin && z <= a.zmax && r <= a.rmax;
    if (inside && !g.loss_code(iz, ir)) {
      st |= ACTIVE;
      atomicAdd(a.new_active, 1ull);
      a.t0[ip] = a.time;
      a.r0[ip] = rold;
      a.z0[ip] = zold;
      if (iz > 0 && g.loss_code(iz - 1, ir))
        adjustFactor = (z - (iz * a.dz + a.zmin)) / (z - zold);
      else
        adjustFactor = (z - a.zmin) / (z - zold);
      if (adjustFactor > 1) adjustFactor = 1;
    }
  }
  if (a.zForcesStart > a.zmin && a.zForcesStart < a.zmax) {
    if (zold < a.zForcesStart) {
      if (z < a.zForcesStart)
        adjustFactor = 0;
      else
        adjustFactor = (z - a.zForcesStart) / (z - zold);
    }
  }
  double pz = a.pz[ip], pr = a.pr[ip], r_pphi = a.rpphi[ip], vphi = a.vphi[ip];
  const double pz0 = pz, pr0 = pr, rp0 = r_pphi;
  if (ip == np - 1) {
    a.last_old[0] = pz0;
    a.last_old[1] = pr0;
    a.last_old[2] = rp0;
  }
  bool lost = false;
  if ((st & ACTIVE) && (z < zmin || z > zmax || r > rmax)) lost = true;
  long long iz = trunc_l((z - zmin) / dz);
  long long ir = trunc_l(fabs(r) / dr);
  if ((st & ACTIVE) && (iz > a.nz || ir > a.nr || iz < 0 || ir < 0)) lost = true;
  // the gather's cells lie within [iz-1, iz+3] x [ir-1, ir+3] of this one
  // (ir is that of |r|, the radius after the reflection below)
  const bool open = iz >= 1 && iz <= a.nz - 3 && ir >= 1 && ir <= a.nr - 3 && a.open[fix(a.ldz, (int)iz, (int)ir)];
  if (!lost && (st & ACTIVE) && !open) {   // an open cell's loss code is 0
    if (g.loss_code(iz, ir)) lost = true;
  }
  bool rneg = false;
  if (r < 0) {
    r *= -1;
    pr *= -1;
    rneg = true;
  }
  const bool track_new = !lost && (st & ACTIVE);
  double Ez, Er, Ephi, Bz, Br, Bphi;
  if (st & ACTIVE) {
    if (open)
      G<true>{a}.interp(z, r, Ez, Er, Ephi, Bz, Br, Bphi);
    else
      g.interp(z, r, Ez, Er, Ephi, Bz, Br, Bphi);
  } else {
    Ez = Er = Ephi = Bz = Br = Bphi = 0;
  }
  // momentum kick, :282-287
  double pphi = (r != 0 ? r_pphi / r : 0);
  double gamma = sqrt(sqr(pr) + sqr(pz) + sqr(pphi) + 1);
  // x / stiffness is x itself when stiffness is 1 (electrons): the division is skipped then
  const bool unit_stiffness = a.stiffness == 1.0;
  auto per_stiffness = [&](double x) { return unit_stiffness ? x : x / a.stiffness; };
  pr += per_stiffness(adjustFactor * dt * ((r != 0 ? sqr(pphi) * a.c / gamma / r : 0) + a.QoMC * (Er + vphi * Bz - vz * Bphi)));
  pz += per_stiffness(adjustFactor * dt * a.QoMC * (Ez + vr * Bphi - vphi * Br));
  r_pphi += per_stiffness(adjustFactor * dt * a.QoMC * r * (Ephi + vz * Br - vr * Bz));
  // velocities, :605-620
  double nvz, nvr, nvphi;
  if (r != 0) {
    gamma = sqrt(sqr(pz) + sqr(pr) + sqr(r_pphi / r) + 1);
    nvphi = a.c * r_pphi / r / gamma;
  } else {
    gamma = sqrt(sqr(pz) + sqr(pr) + 1);
    nvphi = 0;
  }
  nvz = a.c * pz / gamma;
  nvr = a.c * pr / gamma;
  a.z[ip] = z;
  a.r[ip] = r;
  a.pz[ip] = pz;
  a.pr[ip] = pr;
  a.rpphi[ip] = r_pphi;
  a.vz[ip] = nvz;
  a.vr[ip] = nvr;
  a.vphi[ip] = nvphi;
  // extents and status change only now and then: store them only when they do
  {
    const double zmax0 = a.zmax_p[ip], zmin0 = a.zmin_p[ip], rmax0 = a.rmax_p[ip], rmin0 = a.rmin_p[ip];
    double zmx = zmax0, zmn = zmin0, rmx = rmax0, rmn = rmin0;
    if (zold > zmx) zmx = zold;
    if (zold < zmn) zmn = zold;
    if (rold > rmx) rmx = rold;
    if (rold < rmn) rmn = rold;
    if (track_new) {
      if (z > zmx) zmx = z;
      if (z < zmn) zmn = z;
      if (r > rmx) rmx = r;
      if (r < rmn) rmn = r;
    }
    if (zmx != zmax0) a.zmax_p[ip] = zmx;
    if (zmn != zmin0) a.zmin_p[ip] = zmn;
    if (rmx != rmax0) a.rmax_p[ip] = rmx;
    if (rmn != rmin0) a.rmin_p[ip] = rmn;
  }
  if (lost) st = (st & ~ACTIVE) | LOST;
  if (st != st0) a.status[ip] = st;
  // screen crossings, tested as in :370 and :399
  uint32_t ev = (lost ? EV_LOST : 0u) | (rneg ? EV_RNEG : 0u);
  for (int s = 0; s < a.nscreens; s++) {
    double zs = a.sz[s];
    bool fwd = zs <= z && zs > zold && a.sdir[s] == 1;
    bool bwd = !fwd && zs >= z && zs < zold && a.sdir[s] == -1;
    if (fwd || bwd) ev |= 1u << (2 + s);
  }
  bool any = (ev & ~EV_RNEG) != 0;
  if (any) {
    a.evt[ip] = ev;
    a.ev_zold[ip] = zold;
    a.ev_rold[ip] = rold;
    a.ev_vz[ip] = vz;
    a.ev_vr[ip] = vr;
    a.ev_pz0[ip] = pz0;
    a.ev_pr0[ip] = pr0;
    a.ev_rp0[ip] = rp0;
  }
  return any;
}

// one thread per particle; each workgroup also counts its events for the
// ordered event list that rows_kernel builds
__device__ inline void push_block(const PushArgs &a, int64_t n_ub, int bid) {
  const int64_t ip = (int64_t)bid * blockDim.x + threadIdx.x;
  const int64_t np = *a.np;
  bool any = false;
  if (ip < np && ip < n_ub) any = push_one(a, ip, np);
  if (ip < n_ub) a.evflag[ip] = any ? 1 : 0;
  const int cnt = __syncthreads_count(any);
  if (threadIdx.x == 0) {
    a.pcount[bid] = cnt;
    if (cnt) atomicAdd(a.ev_total, (unsigned long long)cnt);
  }
}

// waves per SIMD the push is compiled for (0: the compiler's choice), a build-time constant
#ifndef SPF_PUSH_WAVES
#define SPF_PUSH_WAVES 0
#endif
#if SPF_PUSH_WAVES > 0
#define SPF_PUSH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SPF_PUSH_WAVES, SPF_PUSH_WAVES)))
#else
#define SPF_PUSH_OCCUPANCY
#endif

__global__ __launch_bounds__(256) SPF_PUSH_OCCUPANCY void push_kernel(PushArgs a, int64_t n_ub) {
  push_block(a, n_ub, (int)blockIdx.x);
}

__device__ inline int ragged_run_p(const int32_t *first, int n, int b) {
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

// Ragged push over several runs: workgroups [first[r], first[r+1]) push run r's particles.
struct PushMultiRun {
  PushArgs a;
  int64_t n_ub;
};

__global__ __launch_bounds__(256) void push_multi_kernel(const PushMultiRun *__restrict__ T,
                                                          const int32_t *__restrict__ first, int nruns) {
  const int r = ragged_run_p(first, nruns, (int)blockIdx.x);
  push_block(T[r].a, T[r].n_ub, (int)blockIdx.x - first[r]);
}

// src/secondaryEmission.c:239-372
__device__ void find_loss_coordinates(const RowsArgs &a, double *zReturn, double *rReturn, short *directionCode,
                                      double z0, double r0, double z1, double r1) {
  double zm, rm;
  long long iz1, ir1, iz0, ir0;
  auto FL = [&](long long iz, long long ir) -> int16_t {
    if (iz < 0 || iz > a.nz || ir < 0 || ir > a.nr) return 0;  // the reference reads outside its array here
    return a.flags[fix(a.ldz, (int)iz, (int)ir)];
  };
  *directionCode = 0;
  iz0 = trunc_l((z0 - a.zmin) / a.dz);
  ir0 = trunc_l(r0 / a.dr);
  if (z1 < a.zmin)
    iz1 = -1;
  else
    iz1 = trunc_l((z1 - a.zmin) / a.dz);
  if (r1 < 0)
    ir1 = -1;
  else
    ir1 = trunc_l(r1 / a.dr);
  if (r0 == r1 && z0 == z1) {
    *directionCode = 1;
    *zReturn = z0;
    *rReturn = r0;
    return;
  }
  if (ir0 == ir1) {
    *directionCode = iz0 > iz1 ? 1 : 2;
  } else if (iz0 == iz1) {
    *directionCode = ir0 > ir1 ? 4 : 3;
  } else {
    if (z1 < z0) {
      if (z1 < 0)
        *directionCode = 1;
      else if (ir0 != ir1) {
        if (ir1 >= 0 && ir1 < a.nr && FL(iz0, ir1) & 3)
          *directionCode = 1;
        else if (ir1 < 0 && FL(iz0, -ir1) & 3)
          *directionCode = 1;
      } else if (ir0 == ir1 && (ir0 < (a.nr - 1) && FL(iz0, ir0 + 1) & 3))
        *directionCode = 1;
      else
        *directionCode = r1 < r0 ? 4 : 3;
    } else if (z1 > z0) {
      if (z1 > a.zmax)
        *directionCode = 2;
      else if (ir0 != ir1) {
        if (ir1 >= 0 && ir1 < a.nr && FL(iz1, ir1) & 3)
          *directionCode = 2;
        else if (ir1 < 0 && FL(iz1, -ir1) & 3)
          *directionCode = 2;
      } else if (ir0 == ir1 && (ir0 < (a.nr - 1) && FL(iz1, ir0 + 1) & 3))
        *directionCode = 2;
      else
        *directionCode = r1 < r0 ? 4 : 3;
    } else
      *directionCode = r1 < r0 ? 4 : 3;
  }
  if (*directionCode == 1) {
    zm = ((long long)((z0 - a.zmin) / a.dz)) * a.dz + a.zmin + a.dz * 1e-12;
    rm = (r1 - r0) / (z1 - z0) * (zm - z0) + r0;
  } else if (*directionCode == 2) {
    zm = ((long long)((z0 - a.zmin) / a.dz)) * a.dz + a.zmin + a.dz * (1 - 1e-12);
    rm = (r1 - r0) / (z1 - z0) * (zm - z0) + r0;
  } else if (*directionCode == 4) {
    rm = ((long long)(r0 / a.dr)) * a.dr + a.dr * 1e-12;
    zm = (z1 - z0) / (r1 - r0) * (rm - r0) + z0;
  } else {
    rm = ((long long)(r0 / a.dr)) * a.dr + a.dr * (1 - 1e-12);
    zm = (z1 - z0) / (r1 - r0) * (rm - r0) + z0;
  }
  *zReturn = zm;
  *rReturn = fabs(rm);
}

// screen row of particle ip for screen s, src/advance_particles.c:362-425
__device__ inline void make_screen_row(const RowsArgs &a, int32_t ip, int s, double dt, spf_screen_row &row) {
  double zold = a.ev_zold[ip], vzz = a.ev_vz[ip], vra = a.ev_vr[ip];
  double znew = a.z[ip], rnew = a.r[ip];
  double r0v = a.ev_rold[ip];
  for (int js = 0; js <= s; js++) {
    double zs = a.sz[js];
    bool fwd = zs <= znew && zs > zold && a.sdir[js] == 1;
    bool bwd = !fwd && zs >= znew && zs < zold && a.sdir[js] == -1;
    if (!fwd && !bwd) continue;
    double dtp = (zs - zold) / vzz;
    if (fwd) {
      if (js != s) continue;
      double spz0 = a.compat ? a.last_old[0] : a.ev_pz0[ip];
      double spr0 = a.compat ? a.last_old[1] : a.ev_pr0[ip];
      double srp0 = a.compat ? a.last_old[2] : a.ev_rp0[ip];
      double pphi_a = rnew != 0 ? a.rpphi[ip] / rnew : 0;
      double pphi0 = r0v != 0 ? srp0 / r0v : 0;
      row.r = r0v + vra * dtp;
      row.pz = spz0 + (a.pz[ip] - spz0) / dt * dtp;
      row.pr = spr0 + (a.pr[ip] - spr0) / dt * dtp;
      row.pphi = pphi0 + (pphi_a - pphi0) / dt * dtp;
    } else {
      r0v += vra * dtp;
      if (js != s) continue;
      row.r = r0v;
      row.pz = a.pz[ip];
      row.pr = a.pr[ip];
      row.pphi = r0v != 0 ? a.rpphi[ip] / r0v : 0;
    }
    row.t = (a.time - dt + dtp);
  }
  row.q = a.Q[ip];
  row.r0 = a.r0[ip];
  row.t0 = a.t0[ip];
  row.rmin = a.rmin_p[ip];
  row.rmax = a.rmax_p[ip];
  row.zmin = a.zmin_p[ip];
  row.zmax = a.zmax_p[ip];
  row.ip = ip;
  row.step = a.step;
  row.particleID = a.pid[ip];
  row.generation = a.gen[ip];
  row.screen = (int16_t)s;
}

// lost row of particle ip, src/advance_particles.c:440-480 (columns :581-599)
__device__ inline void make_lost_row(const RowsArgs &a, int32_t ip, spf_lost_row &row) {
  double zLost, rLost;
  short code;
  double rr = a.r[ip] * ((a.evt[ip] & EV_RNEG) ? -1 : 1);
  find_loss_coordinates(a, &zLost, &rLost, &code, a.ev_zold[ip], a.ev_rold[ip], a.z[ip], rr);
  row.t = a.time;
  row.rLost = (float)rLost;
  row.zLost = (float)zLost;
  row.pz = (float)a.pz[ip];
  row.pr = (float)a.pr[ip];
  row.r0 = (float)a.r0[ip];
  row.z0 = (float)a.z0[ip];
  row.q = (float)a.Q[ip];
  row.particleID = a.pid[ip];
  row.ip = ip;
  row.step = a.step;
  row.generation = a.gen[ip];
  row.directionCode = code;
  row.pad = 0;
}

constexpr int ROWS_THREADS = 1024;

// exclusive prefix of a 0/1 flag over the workgroup; returns the total
__device__ int block_flag_scan(bool f, int *wsum, int &excl) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long m = __ballot(f);
  unsigned long long lt = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
  int in_wave = __popcll(lt);
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int before = 0, total = 0;
  for (int k = 0; k < nw; k++) {
    int v = wsum[k];
    if (k < w) before += v;
    total += v;
  }
  __syncthreads();
  excl = before + in_wave;
  return total;
}

// Builds the step's screen rows (grouped per screen, index order within a
// screen) and lost rows (index order), then compacts the particle arrays.
__device__ inline void rows_body(const RowsArgs &a) {
  __shared__ int wsum[ROWS_THREADS / 64];
  __shared__ long long sh_first_lost;
  __shared__ int qblk[ROWS_THREADS];
  __shared__ long long qoff[ROWS_THREADS];
  // ---- ordered event list: push workgroups in order (a scan of their
  // counts), then each workgroup's flagged particles in order (ballots)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = ROWS_THREADS / 64;
  if (*a.ev_total == 0) {
    // no particle crossed a screen or was lost: no rows, nothing to compact;
    // only this step's activations reach the counters
    if (threadIdx.x == 0) {
      *a.n_ev = 0;
      a.np_active[0] += (int64_t)*a.new_active;
      *a.new_active = 0;
    }
    return;
  }
  int64_t n_ev = 0;
  for (int64_t c = 0; c < a.npb; c += blockDim.x) {
    const int64_t b = c + threadIdx.x;
    const int cnt = b < a.npb ? a.pcount[b] : 0;
    // exclusive prefix of the counts, and the list of workgroups with events
    int excl_n, excl_b;
    const int tot_b = block_flag_scan(cnt > 0, wsum, excl_b);
    int v = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      int u = __shfl_up(v, off, 64);
      if (lane >= off) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    int before = 0, total = 0;
    for (int k = 0; k < nwv; k++) {
      before += k < wv ? wsum[k] : 0;
      total += wsum[k];
    }
    __syncthreads();
    excl_n = before + v - cnt;
    if (cnt > 0) {
      qblk[excl_b] = (int)b;
      qoff[excl_b] = n_ev + excl_n;
    }
    __syncthreads();
    // one wavefront per listed workgroup: 256 flags as four ballots
    for (int q = wv; q < tot_b; q += nwv) {
      const int64_t base = (int64_t)qblk[q] * 256;
      int64_t o = qoff[q];
      for (int j = 0; j < 4; j++) {
        const int64_t i = base + 64 * j + lane;
        const bool f = i < a.n_ub && a.evflag[i];
        const unsigned long long m = __ballot(f);
        if (f) a.evlist[o + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
        o += __popcll(m);
      }
    }
    n_ev += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.n_ev = n_ev;
  const int64_t np = a.np[0];
  const double dt = a.dt;
  // ---- screens
  int64_t base = *a.n_srows;
  for (int s = 0; s < a.nscreens; s++) {
    int64_t running = 0;
    for (int64_t c = 0; c < n_ev; c += blockDim.x) {
      int64_t k = c + threadIdx.x;
      int32_t ip = k < n_ev ? a.evlist[k] : 0;
      bool has = k < n_ev && ((a.evt[ip] >> (2 + s)) & 1u);
      int excl;
      int tot = block_flag_scan(has, wsum, excl);
      if (has) {
        spf_screen_row row;
        make_screen_row(a, ip, s, dt, row);
        int64_t pos = base + running + excl;
        if (pos < a.srow_cap)
          a.srows[pos] = row;
        else
          atomicOr(a.overflow, 1);
      }
      running += tot;
    }
    base += running;
  }
  if (threadIdx.x == 0) *a.n_srows = base;
  // ---- lost rows; the ordered lost list goes to move_dst
  int64_t lbase = *a.n_lrows;
  int64_t nlost = 0;
  for (int64_t c = 0; c < n_ev; c += blockDim.x) {
    int64_t k = c + threadIdx.x;
    int32_t ip = k < n_ev ? a.evlist[k] : 0;
    bool has = k < n_ev && (a.evt[ip] & EV_LOST);
    int excl;
    int tot = block_flag_scan(has, wsum, excl);
    if (has) {
      spf_lost_row row;
      make_lost_row(a, ip, row);
      int64_t pos = lbase + nlost + excl;
      if (pos < a.lrow_cap)
        a.lrows[pos] = row;
      else
        atomicOr(a.overflow, 1);
      a.move_dst[nlost + excl] = ip;
    }
    nlost += tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) *a.n_lrows = lbase + nlost;
  // ---- compaction, src/advance_particles.c:490-518
  int64_t new_active = 0;
  if (threadIdx.x == 0) {
    new_active = (int64_t)*a.new_active;
    *a.new_active = 0;   // for the next step's activations
  }
  if (a.compat) {
    // the reference re-tests the same slot without moving its lost flag, so
    // np ends at the first lost index
    if (threadIdx.x == 0) {
      int64_t first = nlost ? a.move_dst[0] : np;
      int64_t removed = np - first;
      a.np[0] = first;
      a.np_active[0] += new_active - removed;
      *a.ev_total = 0;
    }
    return;
  }
  // swap-remove with the flag moved too: the k-th lost slot below the new np
  // receives the k-th survivor counted from the end.  The tail [new_np, np)
  // holds nlost slots; its survivors, taken in descending order, are the
  // sources, one per lost slot below new_np.
  int64_t new_np = np - nlost;
  // m = lost slots below new_np (move_dst is ascending)
  int cnt = 0;
  for (int64_t k = threadIdx.x; k < nlost; k += blockDim.x) cnt += a.move_dst[k] < new_np;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += wsum[k];
    sh_first_lost = t;
  }
  __syncthreads();
  {
    const int64_t m0 = sh_first_lost;
    for (int64_t t = threadIdx.x; t < nlost; t += blockDim.x) a.tflag[t] = 0;
    __syncthreads();
    for (int64_t k = m0 + threadIdx.x; k < nlost; k += blockDim.x) a.tflag[a.move_dst[k] - new_np] = 1;
    __syncthreads();
    // survivor of ascending rank q in the tail is the source of lost slot m0-1-q
    int64_t running = 0;
    for (int64_t c = 0; c < nlost; c += blockDim.x) {
      const int64_t t = c + threadIdx.x;
      const bool surv = t < nlost && !a.tflag[t];
      int excl;
      const int tot = block_flag_scan(surv, wsum, excl);
      if (surv) a.move_src[m0 - 1 - (running + excl)] = (int32_t)(new_np + t);
      running += tot;
    }
  }
  __syncthreads();
  int64_t m = sh_first_lost;
  for (int64_t k = threadIdx.x; k < m; k += blockDim.x) {
    int32_t d = a.move_dst[k], s = a.move_src[k];
    a.Q[d] = a.Q[s];
    a.z[d] = a.z[s];
    a.r[d] = a.r[s];
    a.pz[d] = a.pz[s];
    a.pr[d] = a.pr[s];
    a.vz[d] = a.vz[s];
    a.vr[d] = a.vr[s];
    a.rpphi[d] = a.rpphi[s];
    a.vphi[d] = a.vphi[s];
    a.t0[d] = a.t0[s];
    a.r0[d] = a.r0[s];
    a.z0[d] = a.z0[s];
    a.rmin_p[d] = a.rmin_p[s];
    a.zmin_p[d] = a.zmin_p[s];
    a.rmax_p[d] = a.rmax_p[s];
    a.zmax_p[d] = a.zmax_p[s];
    a.gen[d] = a.gen[s];
    a.pid[d] = a.pid[s];
    a.status[d] = a.status[s];
  }
  if (threadIdx.x == 0) {
    a.np[0] = new_np;
    a.np_active[0] += new_active - nlost;
    *a.ev_total = 0;   // every thread read it before the first barrier
  }
}

__global__ __launch_bounds__(ROWS_THREADS) void rows_kernel(RowsArgs a) { rows_body(a); }

// one workgroup per run: each builds its run's rows and compacts its particles
__global__ __launch_bounds__(ROWS_THREADS) void rows_multi_kernel(const RowsArgs *__restrict__ T) {
  rows_body(T[blockIdx.x]);
}

__global__ void append_kernel(const uint8_t *stage, int64_t n, int64_t cap, double *const *P, int32_t *pid,
                              int16_t *gen, uint32_t *status, const int64_t *np) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t dst = *np + k;
  if (dst >= cap) return;
  const double *d = (const double *)stage;
  for (int j = 0; j < 16; j++) P[j][dst] = d[(int64_t)j * n + k];
  const int32_t *ids = (const int32_t *)(d + 16 * n);
  const int16_t *gens = (const int16_t *)(ids + n);
  const uint32_t *sts = (const uint32_t *)((const uint8_t *)stage + 16 * 8 * n + 4 * n + ((2 * n + 3) / 4) * 4);
  pid[dst] = ids[k];
  gen[dst] = gens[k];
  status[dst] = sts[k];
}

__global__ void append_commit_kernel(int64_t *counters, int64_t n, int64_t n_active, int64_t cap,
                                     int32_t *overflow) {
  int64_t np = counters[0] + n;
  if (np > cap) {
    *overflow |= 2;
    np = cap;
  }
  counters[0] = np;
  counters[1] += n_active;
}

}  // namespace

namespace {

// open[iz, ir] = 1 when every flag in [iz-1, iz+3] x [ir-1, ir+3] has the
// loss-test and gather bits (surface, inside, Ez-zero, Er-zero) clear, for
// 1 <= iz <= nz-3, 1 <= ir <= nr-3; 0 elsewhere
__global__ __launch_bounds__(256) void open_cells_kernel(const int16_t *flags, int ldz, int nz, int nr,
                                                          uint8_t *open) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ldz * (nr + 1)) return;
  const int ir = (int)(q / ldz), iz = (int)(q % ldz);
  uint8_t o = 0;
  if (iz >= 1 && iz <= nz - 3 && ir >= 1 && ir <= nr - 3) {
    constexpr int16_t mask = SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE | SPF_FL_EZ_ZERO | SPF_FL_ER_ZERO;
    int16_t any = 0;
    for (int jr = ir - 1; jr <= ir + 3; jr++)
      for (int jz = iz - 1; jz <= iz + 3; jz++) any |= flags[fix(ldz, jz, jr)];
    o = (any & mask) == 0;
  }
  open[q] = o;
}

}  // namespace

hipError_t launch_open_cells(const int16_t *flags, int ldz, int nz, int nr, uint8_t *open, hipStream_t s) {
  const int64_t n = (int64_t)ldz * (nr + 1);
  hipLaunchKernelGGL(open_cells_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, flags, ldz, nz, nr, open);
  return hipGetLastError();
}

hipError_t launch_push(const PushArgs &a, int64_t n_ub, hipStream_t s) {
  if (n_ub <= 0) return hipSuccess;
  hipLaunchKernelGGL(push_kernel, dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, s, a, n_ub);
  return hipGetLastError();
}

size_t push_multi_table_bytes(int nruns) {
  return sizeof(PushMultiRun) * (size_t)nruns + sizeof(int32_t) * (size_t)(nruns + 1) + sizeof(RowsArgs) * (size_t)nruns +
         64;
}

// the tables of one ragged push: per-run push arguments, first workgroups,
// then (16-byte aligned) the per-run rows arguments
int64_t push_multi_table(const PushArgs *pa, const int64_t *n_ub, const RowsArgs *ra, int nruns, void *host) {
  PushMultiRun *T = (PushMultiRun *)host;
  int32_t *first = (int32_t *)((char *)host + sizeof(PushMultiRun) * (size_t)nruns);
  int64_t total = 0;
  for (int i = 0; i < nruns; i++) {
    T[i].a = pa[i];
    T[i].n_ub = n_ub[i];
    first[i] = (int32_t)total;
    total += (n_ub[i] + 255) / 256;
  }
  first[nruns] = (int32_t)total;
  char *rows = (char *)host + (sizeof(PushMultiRun) * (size_t)nruns + sizeof(int32_t) * (size_t)(nruns + 1) + 15) / 16 * 16;
  memcpy(rows, ra, sizeof(RowsArgs) * (size_t)nruns);
  return total;
}

hipError_t launch_push_multi(const void *dev_table, int nruns, int64_t total_blocks, hipStream_t s) {
  const PushMultiRun *T = (const PushMultiRun *)dev_table;
  const int32_t *first = (const int32_t *)((const char *)dev_table + sizeof(PushMultiRun) * (size_t)nruns);
  const RowsArgs *R = (const RowsArgs *)((const char *)dev_table +
                                         (sizeof(PushMultiRun) * (size_t)nruns + sizeof(int32_t) * (size_t)(nruns + 1) +
                                          15) / 16 * 16);
  if (total_blocks > 0) {
    hipLaunchKernelGGL(push_multi_kernel, dim3((unsigned)total_blocks), dim3(256), 0, s, T, first, nruns);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(rows_multi_kernel, dim3((unsigned)nruns), dim3(ROWS_THREADS), 0, s, R);
  return hipGetLastError();
}

namespace {

// ---- the rows of a large run in several kernels (same rows, same order)
//
// A single workgroup walks a step's events in chunks; with ~10^5 screen
// crossings in a step that takes milliseconds.  Large runs use a pipeline
// with no workgroup waiting for another:
//  1. rows_scan_kernel: exclusive scan of the push workgroups' event counts
//     (or, with no event, only this step's activations into the counters),
//     and the tile counts cleared;
//  2. rows_list_kernel: one wavefront per push workgroup writes its flagged
//     particles, in order, at its offset: the ordered event list;
//     it also counts, per tile of kRowsTile events and row category
//     (screen s, then lost), the events in it;
//  3. rows_tile_scan_kernel: offsets of the tiles per category, and the
//     categories' first rows (screens grouped in screen order);
//  4. rows_write_kernel: every event's rows at (category base + tile offset
//     + rank in the tile), and the ordered lost list;
//  5. rows_compact_kernel: the compaction's sources and destinations
//     (as in rows_body), np and the counters;
//  6. rows_move_kernel: the moves.

__global__ __launch_bounds__(1024) void rows_scan_kernel(RowsArgs a) {
  __shared__ long long ws[16];
  if (*a.ev_total == 0) {
    if (threadIdx.x == 0) {
      *a.n_ev = 0;
      a.np_active[0] += (int64_t)*a.new_active;
      *a.new_active = 0;
    }
    return;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t per = (a.npb + 1023) / 1024;
  const int64_t t0 = std::min<int64_t>((int64_t)threadIdx.x * per, a.npb), t1 = std::min<int64_t>(t0 + per, a.npb);
  long long sum = 0;
  for (int64_t t = t0; t < t1; t++) sum += a.pcount[t];
  long long incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const long long u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) ws[w] = incl;
  __syncthreads();
  long long run = incl - sum;
  for (int k = 0; k < w; k++) run += ws[k];
  for (int64_t t = t0; t < t1; t++) {
    a.evoff[t] = (int32_t)run;
    run += a.pcount[t];
  }
  __shared__ long long sh_nev;
  if (threadIdx.x == 1023) {
    *a.n_ev = run;
    sh_nev = run;
  }
  __syncthreads();
  // the tile counts the list kernel adds to start at 0
  const int64_t ntiles = (sh_nev + kRowsTile - 1) / kRowsTile;
  for (int c = 0; c <= a.nscreens; c++)
    for (int64_t t = threadIdx.x; t < ntiles; t += blockDim.x) a.tcnt[(int64_t)c * a.ntmax + t] = 0;
}

// category c of an event word: screen c (c < nscreens) or lost (c == nscreens)
__device__ inline bool rows_has(const RowsArgs &a, uint32_t e, int c) {
  return c < a.nscreens ? ((e >> (2 + c)) & 1u) != 0 : (e & EV_LOST) != 0;
}

// also counts each tile's events per category (integer additions: the same
// counts in any order), one addition per wavefront chunk, category and tile
__global__ __launch_bounds__(256) void rows_list_kernel(RowsArgs a) {
  if (*a.n_ev == 0) return;
  const int lane = threadIdx.x & 63;
  for (int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); b < a.npb; b += (int64_t)gridDim.x * 4) {
    if (a.pcount[b] == 0) continue;
    int64_t o = a.evoff[b];
    for (int j = 0; j < 4; j++) {
      const int64_t i = b * 256 + 64 * j + lane;
      const bool f = i < a.n_ub && a.evflag[i];
      const unsigned long long m = __ballot(f);
      const int64_t pos = o + __popcll(m & ((1ull << lane) - 1ull));
      const uint32_t e = f ? a.evt[i] : 0u;
      if (f) a.evlist[pos] = (int32_t)i;
      if (m) {
        // the chunk's events fill positions [o, o + popc(m)): at most two tiles;
        // one addition per category and tile from lane 0
        const int64_t t0 = o / kRowsTile, t1 = (o + __popcll(m) - 1) / kRowsTile;
        const unsigned long long in0 = __ballot(f && pos / kRowsTile == t0);
        for (int c = 0; c <= a.nscreens; c++) {
          const unsigned long long hc = __ballot(f && rows_has(a, e, c));
          const int n0 = __popcll(hc & in0), n1 = __popcll(hc) - n0;
          if (lane == 0) {
            if (n0) atomicAdd(a.tcnt + (int64_t)c * a.ntmax + t0, n0);
            if (n1) atomicAdd(a.tcnt + (int64_t)c * a.ntmax + t1, n1);
          }
        }
      }
      o += __popcll(m);
    }
  }
}

__global__ __launch_bounds__(1024) void rows_tile_scan_kernel(RowsArgs a) {
  __shared__ long long ws[16];
  const int64_t nev = *a.n_ev;
  if (nev == 0) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t ntiles = (nev + kRowsTile - 1) / kRowsTile;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t t0 = std::min<int64_t>((int64_t)threadIdx.x * per, ntiles), t1 = std::min<int64_t>(t0 + per, ntiles);
  int64_t sbase = *a.n_srows;
  for (int c = 0; c <= a.nscreens; c++) {
    const int32_t *cnt = a.tcnt + (int64_t)c * a.ntmax;
    int32_t *off = a.toff + (int64_t)c * a.ntmax;
    long long sum = 0;
    for (int64_t t = t0; t < t1; t++) sum += cnt[t];
    long long incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const long long u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    long long run = incl - sum, total = 0;
    for (int k = 0; k < w; k++) run += ws[k];
    for (int k = 0; k < 16; k++) total += ws[k];
    for (int64_t t = t0; t < t1; t++) {
      off[t] = (int32_t)run;
      run += cnt[t];
    }
    if (threadIdx.x == 0) {
      if (c < a.nscreens) {
        a.cbase[c] = sbase;
      } else {
        a.cbase[c] = *a.n_lrows;
        a.cbase[c + 1] = total;   // lost this step
        *a.n_lrows += total;
        *a.n_srows = sbase;
      }
    }
    sbase += total;
    __syncthreads();   // ws is reused by the next category
  }
}

__global__ __launch_bounds__(kRowsTile) void rows_write_kernel(RowsArgs a) {
  __shared__ int wsum[kRowsTile / 64];
  const int64_t nev = *a.n_ev, ntiles = (nev + kRowsTile - 1) / kRowsTile;
  const double dt = a.dt;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t k = tile * kRowsTile + threadIdx.x;
    const int32_t ip = k < nev ? a.evlist[k] : 0;
    const uint32_t e = k < nev ? a.evt[ip] : 0u;
    for (int c = 0; c <= a.nscreens; c++) {
      const bool has = k < nev && rows_has(a, e, c);
      int excl;
      block_flag_scan(has, wsum, excl);
      if (has) {
        const int64_t rank = a.toff[(int64_t)c * a.ntmax + tile] + excl;
        const int64_t pos = a.cbase[c] + rank;
        if (c < a.nscreens) {
          spf_screen_row row;
          make_screen_row(a, ip, c, dt, row);
          if (pos < a.srow_cap)
            a.srows[pos] = row;
          else
            atomicOr(a.overflow, 1);
        } else {
          spf_lost_row row;
          make_lost_row(a, ip, row);
          if (pos < a.lrow_cap)
            a.lrows[pos] = row;
          else
            atomicOr(a.overflow, 1);
          a.move_dst[rank] = ip;
        }
      }
    }
  }
}

// the compaction of the fixed (non-compat) behaviour, as in rows_body
__global__ __launch_bounds__(1024) void rows_compact_kernel(RowsArgs a) {
  __shared__ int wsum[1024 / 64];
  __shared__ long long sh_m0;
  if (*a.n_ev == 0) return;
  const int64_t np = a.np[0], nlost = a.cbase[a.nscreens + 1], new_np = np - nlost;
  int cnt = 0;
  for (int64_t k = threadIdx.x; k < nlost; k += blockDim.x) cnt += a.move_dst[k] < new_np;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int k = 0; k < 16; k++) t += wsum[k];
    sh_m0 = t;
  }
  __syncthreads();
  const int64_t m0 = sh_m0;
  for (int64_t t = threadIdx.x; t < nlost; t += blockDim.x) a.tflag[t] = 0;
  __syncthreads();
  for (int64_t k = m0 + threadIdx.x; k < nlost; k += blockDim.x) a.tflag[a.move_dst[k] - new_np] = 1;
  __syncthreads();
  int64_t running = 0;
  for (int64_t c = 0; c < nlost; c += blockDim.x) {
    const int64_t t = c + threadIdx.x;
    const bool surv = t < nlost && !a.tflag[t];
    int excl;
    const int tot = block_flag_scan(surv, wsum, excl);
    if (surv) a.move_src[m0 - 1 - (running + excl)] = (int32_t)(new_np + t);
    running += tot;
  }
  if (threadIdx.x == 0) {
    a.cbase[a.nscreens + 2] = m0;
    const int64_t new_active = (int64_t)*a.new_active;
    *a.new_active = 0;
    a.np[0] = new_np;
    a.np_active[0] += new_active - nlost;
    *a.ev_total = 0;
  }
}

__global__ __launch_bounds__(256) void rows_move_kernel(RowsArgs a) {
  if (*a.n_ev == 0) return;
  const int64_t m = a.cbase[a.nscreens + 2];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = a.move_dst[k], s = a.move_src[k];
    a.Q[d] = a.Q[s];
    a.z[d] = a.z[s];
    a.r[d] = a.r[s];
    a.pz[d] = a.pz[s];
    a.pr[d] = a.pr[s];
    a.vz[d] = a.vz[s];
    a.vr[d] = a.vr[s];
    a.rpphi[d] = a.rpphi[s];
    a.vphi[d] = a.vphi[s];
    a.t0[d] = a.t0[s];
    a.r0[d] = a.r0[s];
    a.z0[d] = a.z0[s];
    a.rmin_p[d] = a.rmin_p[s];
    a.zmin_p[d] = a.zmin_p[s];
    a.rmax_p[d] = a.rmax_p[s];
    a.zmax_p[d] = a.zmax_p[s];
    a.gen[d] = a.gen[s];
    a.pid[d] = a.pid[s];
    a.status[d] = a.status[s];
  }
}

}  // namespace

hipError_t launch_rows(const RowsArgs &a, hipStream_t s) {
  if (a.compat || a.n_ub < kRowsPipelineMin) {
    hipLaunchKernelGGL(rows_kernel, dim3(1), dim3(ROWS_THREADS), 0, s, a);
    return hipGetLastError();
  }
  const unsigned tiles = (unsigned)std::min<int64_t>((a.n_ub + kRowsTile - 1) / kRowsTile, 1024);
  hipLaunchKernelGGL(rows_scan_kernel, dim3(1), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(rows_list_kernel, dim3((unsigned)std::min<int64_t>((a.npb + 3) / 4, 4096)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(rows_tile_scan_kernel, dim3(1), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(rows_write_kernel, dim3(tiles), dim3(kRowsTile), 0, s, a);
  hipLaunchKernelGGL(rows_compact_kernel, dim3(1), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(rows_move_kernel, dim3((unsigned)std::min<int64_t>((a.n_ub + 255) / 256, 2048)), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

// stage layout: 16 double columns of n, int32 ids, int16 generations (padded
// to 4 bytes), uint32 status
size_t append_stage_bytes(int64_t n) { return 16 * 8 * n + 4 * n + ((2 * n + 3) / 4) * 4 + 4 * n; }

hipError_t launch_append_kernels(Run &R, const void *stage, int64_t n, int64_t n_active, double *const *Pdev,
                                 hipStream_t s) {
  hipLaunchKernelGGL(append_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const uint8_t *)stage, n,
                     R.cap, Pdev, R.pid, R.gen, R.status, R.counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(append_commit_kernel, dim3(1), dim3(1), 0, s, R.counters, n, n_active, R.cap, R.overflow);
  return hipGetLastError();
}

hipError_t launch_append_commit(Run &R, int64_t n, int64_t n_active, hipStream_t s) {
  hipLaunchKernelGGL(append_commit_kernel, dim3(1), dim3(1), 0, s, R.counters, n, n_active, R.cap, R.overflow);
  return hipGetLastError();
}

}  // namespace spf
