// spf_emit.cpp -- host side of cathode emission.  Follows src/cathode.c; the
// per-particle state that does not need a random draw (momenta, the
// spread-over-dt kick and the field reads it makes) is written on the device.
// Build with -ffp-contract=off and without sin/cos builtins so each libm call
// and each FP64 operation is the one written here.
#include "spf_emit.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <time.h>

namespace spf {

static const double kPi = 3.141592653589793238462643;

// ---- uniform and Gaussian draws ----
void philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    if (round) {
      k0 += 0x9E3779B97F4A7C15ull;   // Weyl increments of the key schedule
      k1 += 0xBB67AE8584CAA73Bull;
    }
    unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * c0;
    unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * c2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void EmitRng::seed(uint64_t s) {
  key0 = s;
  block = 0;
  pos = 4;
  draws = 0;
}

uint64_t EmitRng::next64() {
  if (pos == 4) {
    block++;
    uint64_t ctr[4] = {block, 0, 0, 0}, key[2] = {key0, 0};
    philox4x64_10(ctr, key, buf);
    pos = 0;
  }
  draws++;
  return buf[pos++];
}

double EmitRng::uniform() { return (double)(next64() >> 11) * (1.0 / 9007199254740992.0); }

double EmitRng::gauss() {
  if (has_saved) {
    has_saved = false;
    return saved;
  }
  double v1, v2, s;
  do {
    v1 = 2.0 * uniform() - 1.0;
    v2 = 2.0 * uniform() - 1.0;
    s = v1 * v1 + v2 * v2;
  } while (s >= 1.0 || s == 0.0);
  double fac = sqrt(-2.0 * log(s) / s);
  saved = v1 * fac;
  has_saved = true;
  return v2 * fac;
}

double EmitRng::gauss_lim(double mean, double sigma, double limit) {
  double g = gauss();
  if (limit > 0)
    while (fabs(g) > limit) g = gauss();
  return mean + sigma * g;
}

double radical_inverse(uint64_t k, int b) {
  double inv = 1.0 / b, f = inv, x = 0;
  while (k) {
    x += (double)(k % (uint64_t)b) * f;
    k /= (uint64_t)b;
    f *= inv;
  }
  return x;
}

// Linear interpolation of f(x) at xo for x increasing and x[0] <= xo <=
// x[n-1] (the only case src/cathode.c:351-353 calls it in): the bracket is
// found by bisection, with xo equal to a node taking the node's right interval.
static double interp_linear(const std::vector<double> &f, const std::vector<double> &x, double xo) {
  int64_t n = (int64_t)x.size();
  if (n == 1) return f[0];
  int64_t lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (xo < x[mid])
      hi = mid;
    else
      lo = mid;
  }
  return f[lo] + (f[hi] - f[lo]) / (x[hi] - x[lo]) * (xo - x[lo]);
}

static double sqr(double x) { return x * x; }

// ---- define_cathode ----
int cathode_define(const spf_cathode_desc &d, const CathodeEnv &env, EmitRng &rng, std::vector<Cathode> &out,
                   std::string &err) {
  // checks in the order of src/cathode.c:44-84
  const double c_mks = env.c_mks;
  double temperature = d.temperature;
  if (d.z_position < env.zmin || d.z_position > env.zmax) return err = "cathode is outside of simulation region", 1;
  if (d.inner_radius > d.outer_radius) return err = "cathode has negative radial extent", 1;
  if (d.inner_radius == d.outer_radius && d.current_density != 0)
    return err = "cathode has zero radial extent--specify electrons_per_macroparticle instead of current_density", 1;
  if (d.inner_radius < 0 || d.outer_radius > env.rmax)
    return err = "all or part of cathode is outside of simulation region", 1;
  if (!d.field_emission && d.current_density <= 0 && d.electrons_per_macroparticle == 0 && temperature <= 0)
    return err = "either field_emission, current_density, electrons_per_macroparticle, or temperature must be "
                 "non-zero",
           1;
  if (d.field_emission && (d.determine_temperature || d.current_density || d.n_profile))
    return err = "field_emission is incompatible with determine_temperature, current_density, and profile", 1;
  if (d.field_emission && !d.electrons_per_macroparticle)
    return err = "specify electrons_per_macroparticle with field_emission", 1;
  if (d.field_emission && !d.work_function) return err = "specify work_function with field_emission", 1;
  if (d.current_density < 0) return err = "the current density should be a positive quantity", 1;
  if (d.electrons_per_macroparticle < 0) return err = "electrons_per_macroparticle should be a positive quantity", 1;
  if (temperature < 0) return err = "temperature should be a positive quantity", 1;
  if (temperature > 0 && d.work_function > 0 && d.distribution_correction_interval)
    return err = "can't do distribution correction with Richardson-Schottky emission model", 1;
  if (d.stop_time && d.start_time > d.stop_time) return err = "stop_time is non-zero and less than start_time", 1;
  if (d.distribution_correction_interval < 0) return err = "distribution_correction_interval < 0", 1;
  if (d.distribution_correction_interval && d.discretize_radii)
    return err = "distribution correction and discrete radii are incompatible", 1;
  if (d.stiffness <= 0) return err = "stiffness <= 0", 1;
  if (fabs(d.initial_omega * d.outer_radius) >= c_mks) return err = "initial_omega*outer_radius >= c---unphysical", 1;
  if (d.add_thermal_velocities && !(temperature > 0 || d.determine_temperature))
    return err = "add_thermal_velocities is nonzero but the temperature is not positive", 1;
  if (d.n_profile < 0 || (d.n_profile > 0 && (!d.profile_time || !d.profile_factor)))
    return err = "profile arrays missing", 1;
  if (d.e_mks <= 0 || d.me_mks <= 0 || d.k_boltzmann_mks <= 0)
    return err = "physical constants e_mks, me_mks and k_boltzmann_mks must be set", 1;
  // profile and Halton checks, made per cathode in the reference; they do not
  // depend on the cathode, so they come before anything is appended
  if (d.n_profile) {
    if (d.distribution_correction_interval != 0)
      return err = "can't do distribution correction with an emission profile", 1;
    for (int64_t i = 1; i < d.n_profile; i++)
      if (d.profile_time[i - 1] >= d.profile_time[i])
        return err = "profile data is not ordered by increasing time values", 1;
  }
  if (d.halton_radix_dt) {
    if (!d.spread_over_dt) return err = "you have nonzero halton_radix_dt but didn't ask to spread_over_dt", 1;
    if (d.distribution_correction_interval)
      return err = "you can't use distribution_correction_interval with halton sequence", 1;
    if (d.halton_radix_dt < 2) return err = "halton_radix_dt<2", 1;
  }
  if (d.halton_radix_r) {
    if (d.distribution_correction_interval)
      return err = "you can't use distribution_correction_interval with halton sequence", 1;
    if (d.zoned_emission) return err = "you can't use zoned_emission with halton sequence", 1;
    if (d.halton_radix_r < 2) return err = "halton_radix_r<2", 1;
  }

  // src/cathode.c:85-91: a zero seed takes an odd clock value
  int64_t seed = d.random_number_seed;
  if (seed == 0) {
    seed = (int64_t)time(NULL);
    seed = 2 * (seed / 2) + 1;
  }
  rng.seed((uint64_t)(seed < 0 ? -seed : seed));

  // field emission: one sub-cathode per radial cell, src/cathode.c:93-100
  int64_t n_new = 1;
  if (d.field_emission) {
    if (d.inner_radius != 0)
      n_new = (int64_t)((d.outer_radius - d.inner_radius) / env.dr);
    else
      n_new = (int64_t)((d.outer_radius - d.inner_radius + env.dr / 2) / env.dr);
  }

  for (int64_t i = 0; i < n_new; i++) {
    // determine_temperature: Richardson's law solved for T, src/cathode.c:127-141
    if (d.determine_temperature) {
      if (d.work_function <= 0 || d.current_density <= 0)
        return err = "to determine temperature, must have positive work function and current density", 1;
      double lastTemp = 1273;
      for (int it = 0; it < 100; it++) {
        temperature =
            -d.e_mks * d.work_function / d.k_boltzmann_mks / log(d.current_density / 60e4 / sqr(lastTemp));
        if (fabs(lastTemp - temperature) < 0.01) break;
        lastTemp = temperature;
      }
    }
    Cathode k;
    k.sigmaV = 0;
    if (temperature && d.add_thermal_velocities) k.sigmaV = sqrt(temperature * d.k_boltzmann_mks / d.me_mks);
    if (d.n_profile) {
      k.t_prof.assign(d.profile_time, d.profile_time + d.n_profile);
      k.A_prof.assign(d.profile_factor, d.profile_factor + d.n_profile);
    }
    if (!d.field_emission) {
      k.inner_radius = d.inner_radius;
      k.outer_radius = d.outer_radius;
    } else {
      k.outer_radius = (i + 0.5) * env.dr;
      if ((k.inner_radius = k.outer_radius - env.dr) < 0) k.inner_radius = 0;
      if (i == n_new - 1) k.outer_radius = d.outer_radius;
      if (k.outer_radius > d.outer_radius) k.outer_radius = d.outer_radius;
      if (k.inner_radius < d.inner_radius) k.inner_radius = d.inner_radius;
    }
    k.z_position = d.z_position;
    k.initial_vz = d.initial_pz / sqrt(1 + sqr(d.initial_pz)) * c_mks;
    k.initial_omega = d.initial_omega;
    k.current_density = d.current_density;
    k.temperature = temperature;
    k.work_function = d.work_function;
    k.field_emission = d.field_emission;
    k.field_emission_beta = d.field_emission_beta;
    k.start_time = d.start_time;
    k.stop_time = d.stop_time;
    k.time_offset = d.time_offset;
    k.autophase = d.autophase;
    k.number_per_step = d.number_per_step;
    k.QoMC = -d.e_mks / (d.me_mks * c_mks);
    k.correction_interval = d.distribution_correction_interval;
    k.start_step = -1;
    k.step_emission_started = -1;
    k.stiffness = d.stiffness;
    k.electrons_per_macroparticle = d.electrons_per_macroparticle;
    k.spread_over_dt = d.spread_over_dt;
    k.zoned_emission = d.zoned_emission;
    if ((k.discretize_radii = d.discretize_radii)) k.correction_interval = 0;
    if (k.field_emission) k.zoned_emission = 0;
    if (!k.zoned_emission) k.correction_interval = 0;
    k.halton_radix_dt = d.halton_radix_dt;
    k.halton_radix_r = d.halton_radix_r;
    k.e_mks = d.e_mks;
    k.me_mks = d.me_mks;
    k.k_boltzmann_mks = d.k_boltzmann_mks;
    out.push_back(k);
  }
  return 0;
}

// ---- current-density models ----
double field_emission_current_density(double E, double beta, double wf) {
  E = fabs(E) * beta;
  double y = 3.79e-5 * sqrt(E) / wf;
  double t = ((-0.0020 * y + 0.0131) * y + 0.1091) * y + 0.9899;
  double v = ((0.0445 * y - 0.6782) * y - 0.4091) * y + 1.0437;
  return 1.54e-6 * sqr(E) / (wf * sqr(t)) * exp(-6.83e9 * v * pow(wf, 1.5) / E);
}

bool cathode_needs_ez(const Cathode &k) { return (k.temperature > 0 && k.work_function) || k.field_emission; }

void cathode_ez_index(const Cathode &k, const CathodeEnv &env, int64_t *iz, int64_t *ir) {
  int64_t z = (int64_t)((k.z_position - env.zmin) / env.dz + 1.5);
  int64_t r = (int64_t)((k.inner_radius + k.outer_radius) / 2.0 / env.dr);
  if (z <= 0) z = 1;
  *iz = z;
  *ir = r;
}

void cathode_set_current_density(Cathode &k, double Ez) {
  if (Ez >= 0 && k.initial_vz <= 0) return;
  if (k.temperature > 0 && k.work_function) {
    // Richardson-Schottky, src/cathode.c:697-709
    double wfCorrection = 0.012 * sqrt(fabs(Ez * 1e-5));
    if (wfCorrection > k.work_function) wfCorrection = k.work_function;
    k.current_density = 60e4 * sqr(k.temperature) *
                        exp(-(k.work_function - wfCorrection) / (k.temperature * k.k_boltzmann_mks / k.e_mks));
  } else if (k.field_emission) {
    k.current_density = field_emission_current_density(Ez, k.field_emission_beta, k.work_function);
  }
}

// addThermalVelocities, src/cathode.c:761-772: the kicks added to vr, vphi, vz
static void thermal_kick(EmitRng &rng, double sigmaV, double *dvr, double *dvphi, double *dvz) {
  double vx = rng.gauss_lim(0.0, sigmaV, 0.0);
  double vy = rng.gauss_lim(0.0, sigmaV, 0.0);
  double theta = rng.uniform() * (2 * kPi);
  double sin_theta = sin(theta);
  double cos_theta = cos(theta);
  *dvr = vx * cos_theta + vy * sin_theta;
  *dvphi = -vx * sin_theta + vy * cos_theta;
  *dvz = fabs(rng.gauss_lim(0.0, sigmaV, 0.0));
}

// ---- emit_electrons_from_cathode ----
int cathode_emit(Cathode &k, const CathodeEnv &env, EmitRng &rng, EmitBatch &out, std::string &err) {
  out.n = 0;
  out.thermal = k.sigmaV != 0;
  out.spread = k.spread_over_dt != 0;
  out.cols.clear();
  k.last_count = 0;
  if (k.number_per_step == 0 && k.electrons_per_macroparticle == 0) return 0;

  if (k.start_step == -1) {
    // first call: rates, step window and radial bins, src/cathode.c:271-345
    if (k.current_density <= 1. / sqrt(DBL_MAX)) return 0;
    k.ref_current_density = k.current_density;
    if (k.autophase) {
      k.stop_time = (k.stop_time - k.start_time) + env.time + k.time_offset;
      k.start_time = env.time + k.time_offset;
    }
    if (env.dti <= 0) return err = "integration step size is not defined (emit_electrons_from_cathode)", 1;
    k.inner_r2 = sqr(k.inner_radius);
    k.outer_r2 = sqr(k.outer_radius);
    double area;
    if ((area = kPi * (k.outer_r2 - k.inner_r2)) != 0) {
      if (k.electrons_per_macroparticle) {
        k.number_per_step = area * env.dti * k.current_density / (k.e_mks * k.electrons_per_macroparticle);
        if (k.number_per_step == 0) return 0;
      } else {
        k.electrons_per_macroparticle = area * env.dti * k.current_density / (k.e_mks * k.number_per_step);
      }
    } else {
      k.correction_interval = 0;
      k.n_bins = 1;
    }
    k.start_step = (int64_t)((k.start_time - env.start_time) / env.dti);
    if (k.stop_time && k.stop_time > env.start_time)
      k.stop_step = (int64_t)((k.stop_time - env.start_time) / env.dti);
    else
      k.stop_step = LONG_MAX;
    k.Q = -k.e_mks * k.electrons_per_macroparticle;
    k.n_bins = (int64_t)((k.outer_radius - k.inner_radius) / env.dr + 1);
    k.distribution.assign(k.n_bins, 0);
    k.deficit.assign(k.n_bins, 0);
    k.area.assign(k.n_bins, 0.0);
    k.rmin.assign(k.n_bins, 0.0);
    k.n_emitted = 0;
    k.totalArea = 0;
    for (int64_t ib = 0; ib < k.n_bins; ib++) {
      int64_t ir = (int64_t)(k.inner_radius / env.dr + ib);
      if (ir == 0) {
        double r_max = env.dr / 2;
        k.area[ib] = sqr(r_max) * kPi;
        k.rmin[ib] = 0;
        // the axis bin's area is left out of totalArea, as in the reference
      } else {
        double r_min = env.dr * (ir - 0.5);
        double r_max = env.dr * (ir + 0.5);
        if ((k.area[ib] = kPi * (sqr(r_max) - sqr(r_min))) < 0) return err = "negative area found for cathode annulus", 1;
        k.rmin[ib] = r_min;
        k.totalArea += k.area[ib];
      }
    }
  }

  double profile_factor = 1;
  if (!k.t_prof.empty()) {
    if (env.time >= k.t_prof.front() && env.time <= k.t_prof.back())
      profile_factor = interp_linear(k.A_prof, k.t_prof, env.time);
    else
      profile_factor = 0;
  }

  if (k.start_step > env.time_step || k.stop_step < env.time_step) return 0;
  if (k.step_emission_started < 0) k.step_emission_started = env.time_step;

  double dr2 = (k.outer_r2 - k.inner_r2) / k.number_per_step;
  double total_area = k.totalArea;
  double totalNominal = k.number_per_step * (k.current_density / k.ref_current_density);
  double total = totalNominal;
  if (k.correction_interval && env.time_step % k.correction_interval == 0) {
    // bring every annulus up to its share of everything emitted so far, src/cathode.c:382-399
    total = 0;
    for (int64_t ib = 0; ib < k.n_bins; ib++) {
      double want = (env.time_step - k.step_emission_started + 1) * (k.area[ib] / total_area) * totalNominal + 0.5;
      if ((k.deficit[ib] = (int64_t)(want - k.distribution[ib])) <= 0) k.deficit[ib] = 0;
      total += k.deficit[ib];
    }
  } else if (k.zoned_emission) {
    // per-annulus random rounding, src/cathode.c:400-422
    total = 0;
    for (int64_t ib = 0; ib < k.n_bins; ib++) {
      if (total_area) {
        double fnumber = profile_factor * (k.area[ib] / total_area) * totalNominal;
        if (fnumber > 1) {
          k.deficit[ib] = (int64_t)fnumber;
          fnumber -= k.deficit[ib];
        } else {
          k.deficit[ib] = 0;
        }
        if (fnumber && fnumber > rng.uniform()) k.deficit[ib] += 1;
        total += k.deficit[ib];
      }
    }
  } else {
    double fnumber = total * profile_factor;
    total = (double)(int64_t)fnumber;
    fnumber -= (double)(int64_t)fnumber;
    if (fnumber && fnumber > rng.uniform()) total += 1;
  }
  if (total < 0) return err = "negative number of particles created!", 1;
  if (total > 9.0e15) return err = "integer overflow in computing particle array dimension", 1;
  if (!total) return 0;

  int64_t n = (int64_t)total;
  out.n = n;
  out.cols.assign((size_t)5 * n, 0.0);
  double *R = out.cols.data(), *DVR = R + n, *DVPHI = R + 2 * n, *DVZ = R + 3 * n, *DTE = R + 4 * n;
  if (out.spread) {
    int64_t iz = (int64_t)((k.z_position - env.zmin) / env.dz);
    if (iz <= 0) iz = 1;
    if (iz > env.nz) iz = env.nz;
    out.iz_spread = (int)iz;
  }
  auto kick = [&](int64_t j) {
    if (k.sigmaV) thermal_kick(rng, k.sigmaV, DVR + j, DVPHI + j, DVZ + j);
  };

  int64_t j = 0;
  if (k.correction_interval) {
    // annulus by annulus, src/cathode.c:458-525
    for (int64_t ib = 0; ib < k.n_bins; ib++) {
      for (int64_t q = 0; q < k.deficit[ib]; q++, j++) {
        R[j] = sqrt(rng.uniform() * k.area[ib] / kPi + sqr(k.rmin[ib]));
        kick(j);
        if (out.spread) DTE[j] = env.dti * q / (1.0 * k.deficit[ib]);
      }
      k.distribution[ib] += k.deficit[ib];
    }
    if (j != n) return err = "wrong number of particles emitted for zoned cathode", 1;
  } else if (!k.discretize_radii) {
    // uniform in area over the cathode, src/cathode.c:526-597
    for (; j < n; j++) {
      double rn;
      if (k.halton_radix_r)
        rn = radical_inverse(++k.halton_k_r, k.halton_radix_r);
      else
        rn = rng.uniform();
      R[j] = sqrt(rn * (sqr(k.outer_radius) - sqr(k.inner_radius)) + sqr(k.inner_radius));
      kick(j);
      if (out.spread) {
        if (!k.halton_radix_dt)
          DTE[j] = env.dti * j / (1.0 * total);
        else
          DTE[j] = env.dti * radical_inverse(++k.halton_k_dt, k.halton_radix_dt);
      }
    }
  } else {
    // equal-area radii, src/cathode.c:598-656
    for (; j < n; j++) {
      R[j] = sqrt(k.inner_r2 + (j + 1) * dr2);
      kick(j);
      if (out.spread) DTE[j] = env.dti * j / (1.0 * n);
    }
  }
  k.n_emitted += n;
  k.last_count = n;
  return 0;
}

}  // namespace spf
