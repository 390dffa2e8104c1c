/*
 * spf_oracle_cathode.c -- CPU restatement of spiffe's cathode emission
 * (src/cathode.c), with the build's own random stream in place of the SDDS
 * toolkit's random_1 / gauss_rn_lim / Halton routines.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's emission path; only
 * tests/, the smoke check and bench.py's CPU-baseline leg load it.
 *
 * The stream: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by
 * {|seed|, 0}; counter block b = 1, 2, ... and its four 64-bit words in order;
 * a uniform double is (word >> 11) * 2^-53.  Gaussian deviates come from the
 * Marsaglia polar method, the second deviate of each pair served by the next
 * call.  Halton points are van der Corput radical inverses of k = 1, 2, ...
 * Build with -ffp-contract=off and without sin/cos builtins.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "spf_oracle.h"

#define O_PI 3.141592653589793238462643

static double osqr(double x) { return x * x; }

/* ---- the stream ---- */
void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t x[4], k[2];
  int i;
  memcpy(x, ctr, sizeof(x));
  memcpy(k, key, sizeof(k));
  for (i = 0; i < 10; i++) {
    unsigned __int128 a, b;
    uint64_t y[4];
    if (i > 0) {
      k[0] += 0x9E3779B97F4A7C15ull;
      k[1] += 0xBB67AE8584CAA73Bull;
    }
    a = (unsigned __int128)x[0] * 0xD2E7470EE14C6C93ull;
    b = (unsigned __int128)x[2] * 0xCA5A826395121157ull;
    y[0] = (uint64_t)(b >> 64) ^ x[1] ^ k[0];
    y[1] = (uint64_t)b;
    y[2] = (uint64_t)(a >> 64) ^ x[3] ^ k[1];
    y[3] = (uint64_t)a;
    memcpy(x, y, sizeof(x));
  }
  memcpy(out, x, sizeof(x));
}

void oracle_rng_seed(orng *g, uint64_t seed) {
  g->key0 = seed;
  g->block = 0;
  g->pos = 4;
}

uint64_t oracle_rng_next64(orng *g) {
  if (g->pos >= 4) {
    uint64_t ctr[4], key[2];
    g->block += 1;
    ctr[0] = g->block;
    ctr[1] = ctr[2] = ctr[3] = 0;
    key[0] = g->key0;
    key[1] = 0;
    oracle_philox4x64_10(ctr, key, g->buf);
    g->pos = 0;
  }
  return g->buf[g->pos++];
}

double oracle_rng_uniform(orng *g) { return (double)(oracle_rng_next64(g) >> 11) / 9007199254740992.0; }

double oracle_rng_gauss(orng *g) {
  double u, v, s, m;
  if (g->has_saved) {
    g->has_saved = 0;
    return g->saved;
  }
  for (;;) {
    u = 2.0 * oracle_rng_uniform(g) - 1.0;
    v = 2.0 * oracle_rng_uniform(g) - 1.0;
    s = u * u + v * v;
    if (s < 1.0 && s != 0.0)
      break;
  }
  m = sqrt(-2.0 * log(s) / s);
  g->saved = u * m;
  g->has_saved = 1;
  return v * m;
}

/* gauss_rn_lim(mean, sigma, limit, random_1) as used by src/cathode.c:764-771 */
static double gauss_rn_lim(orng *g, double mean, double sigma, double limit) {
  double x;
  do
    x = oracle_rng_gauss(g);
  while (limit > 0 && fabs(x) > limit);
  return mean + sigma * x;
}

double oracle_radical_inverse(uint64_t k, int32_t base) {
  double scale = 1.0 / base, digit_weight = scale, value = 0;
  for (; k > 0; k /= (uint64_t)base) {
    value += (double)(k % (uint64_t)base) * digit_weight;
    digit_weight *= scale;
  }
  return value;
}

/* ---- the cathode set ---- */
typedef struct {
  double z_position, initial_vz, initial_omega, inner_radius, outer_radius;
  double current_density, ref_current_density, temperature, work_function;
  int32_t field_emission;
  double field_emission_beta, start_time, stop_time, time_offset;
  int64_t start_step, stop_step;
  int32_t autophase;
  double number_per_step;
  int64_t step_emission_started;
  double QoMC, Q, electrons_per_macroparticle, inner_r2, outer_r2, stiffness;
  int32_t discretize_radii;
  int64_t correction_interval;
  int32_t zoned_emission;
  int64_t n_bins, n_emitted, last_count;
  int32_t spread_over_dt;
  int64_t *distribution, *deficit;
  double *area, *rmin, totalArea;
  int32_t halton_radix_dt, halton_radix_r;
  uint64_t halton_n_dt, halton_n_r;
  int64_t n_prof;
  double *t_prof, *A_prof;
  double sigmaVMaxwellian;
  double e_mks, me_mks, k_boltzmann_mks;
} ocathode;

struct ocathode_set {
  ocathode *k;
  int32_t n;
  orng rng;
  int64_t next_id;
};

ocathode_set *oracle_cathodes_create(void) {
  ocathode_set *s = (ocathode_set *)calloc(1, sizeof(*s));
  s->next_id = 1;
  s->rng.pos = 4;
  return s;
}

void oracle_cathodes_destroy(ocathode_set *s) {
  int32_t i;
  if (!s)
    return;
  for (i = 0; i < s->n; i++) {
    free(s->k[i].distribution);
    free(s->k[i].deficit);
    free(s->k[i].area);
    free(s->k[i].rmin);
    free(s->k[i].t_prof);
    free(s->k[i].A_prof);
  }
  free(s->k);
  free(s);
}

orng *oracle_cathodes_rng(ocathode_set *s) { return &s->rng; }
int64_t oracle_cathodes_next_id(ocathode_set *s) { return s->next_id; }

static int bomb(char *err, int64_t errlen, const char *msg) {
  if (err && errlen > 0)
    snprintf(err, (size_t)errlen, "%s", msg);
  return 1;
}

/* process_cathode_definition, src/cathode.c:21-237 */
int oracle_cathode_define(ocathode_set *s, const ofields *f, const ocathode_desc *d, int32_t *first,
                          int32_t *count, char *err, int64_t errlen) {
  double temperature = d->temperature;
  int64_t newCathodes, i, seed;
  if (d->z_position < f->zmin || d->z_position > f->zmax)
    return bomb(err, errlen, "cathode is outside of simulation region");
  if (d->inner_radius > d->outer_radius)
    return bomb(err, errlen, "cathode has negative radial extent");
  if (d->inner_radius == d->outer_radius && d->current_density != 0)
    return bomb(err, errlen,
                "cathode has zero radial extent--specify electrons_per_macroparticle instead of current_density");
  if (d->inner_radius < 0 || d->outer_radius > f->rmax)
    return bomb(err, errlen, "all or part of cathode is outside of simulation region");
  if (!d->field_emission && d->current_density <= 0 && d->electrons_per_macroparticle == 0 && temperature <= 0)
    return bomb(err, errlen,
                "either field_emission, current_density, electrons_per_macroparticle, or temperature must be "
                "non-zero");
  if (d->field_emission && (d->determine_temperature || d->current_density || d->n_profile))
    return bomb(err, errlen, "field_emission is incompatible with determine_temperature, current_density, and profile");
  if (d->field_emission && !d->electrons_per_macroparticle)
    return bomb(err, errlen, "specify electrons_per_macroparticle with field_emission");
  if (d->field_emission && !d->work_function)
    return bomb(err, errlen, "specify work_function with field_emission");
  if (d->current_density < 0)
    return bomb(err, errlen, "the current density should be a positive quantity");
  if (d->electrons_per_macroparticle < 0)
    return bomb(err, errlen, "electrons_per_macroparticle should be a positive quantity");
  if (temperature < 0)
    return bomb(err, errlen, "temperature should be a positive quantity");
  if (temperature > 0 && d->work_function > 0 && d->distribution_correction_interval)
    return bomb(err, errlen, "can't do distribution correction with Richardson-Schottky emission model");
  if (d->stop_time && d->start_time > d->stop_time)
    return bomb(err, errlen, "stop_time is non-zero and less than start_time");
  if (d->distribution_correction_interval < 0)
    return bomb(err, errlen, "distribution_correction_interval < 0");
  if (d->distribution_correction_interval && d->discretize_radii)
    return bomb(err, errlen, "distribution correction and discrete radii are incompatible");
  if (d->stiffness <= 0)
    return bomb(err, errlen, "stiffness <= 0");
  if (fabs(d->initial_omega * d->outer_radius) >= f->c_mks)
    return bomb(err, errlen, "initial_omega*outer_radius >= c---unphysical");
  if (d->add_thermal_velocities && !(temperature > 0 || d->determine_temperature))
    return bomb(err, errlen, "add_thermal_velocities is nonzero but the temperature is not positive");

  seed = d->random_number_seed;
  if (seed == 0) {
    seed = (int64_t)time(NULL);
    seed = 2 * (seed / 2) + 1;
  }
  oracle_rng_seed(&s->rng, (uint64_t)llabs(seed));

  newCathodes = 1;
  if (d->field_emission) {
    if (d->inner_radius != 0)
      newCathodes = (int64_t)((d->outer_radius - d->inner_radius) / f->dr);
    else
      newCathodes = (int64_t)((d->outer_radius - d->inner_radius + f->dr / 2) / f->dr);
  }
  s->k = (ocathode *)realloc(s->k, sizeof(ocathode) * (size_t)(s->n + newCathodes + 1));
  if (first)
    *first = s->n;
  for (i = 0; i < newCathodes; i++) {
    ocathode *k = s->k + s->n;
    memset(k, 0, sizeof(*k));
    if (d->determine_temperature) {
      double lastTemp = 1273;
      int it;
      if (d->work_function <= 0 || d->current_density <= 0)
        return bomb(err, errlen, "to determine temperature, must have positive work function and current density");
      for (it = 0; it < 100; it++) {
        temperature = -d->e_mks * d->work_function / d->k_boltzmann_mks /
                      log(d->current_density / 60e4 / osqr(lastTemp));
        if (fabs(lastTemp - temperature) < 0.01)
          break;
        lastTemp = temperature;
      }
    }
    k->sigmaVMaxwellian = 0;
    if (temperature && d->add_thermal_velocities)
      k->sigmaVMaxwellian = sqrt(temperature * d->k_boltzmann_mks / d->me_mks);
    if (d->n_profile) {
      int64_t j;
      if (d->distribution_correction_interval != 0)
        return bomb(err, errlen, "can't do distribution correction with an emission profile");
      for (j = 1; j < d->n_profile; j++)
        if (d->profile_time[j - 1] >= d->profile_time[j])
          return bomb(err, errlen, "profile data is not ordered by increasing time values");
      k->n_prof = d->n_profile;
      k->t_prof = (double *)malloc(sizeof(double) * (size_t)d->n_profile);
      k->A_prof = (double *)malloc(sizeof(double) * (size_t)d->n_profile);
      memcpy(k->t_prof, d->profile_time, sizeof(double) * (size_t)d->n_profile);
      memcpy(k->A_prof, d->profile_factor, sizeof(double) * (size_t)d->n_profile);
    }
    if (!d->field_emission) {
      k->inner_radius = d->inner_radius;
      k->outer_radius = d->outer_radius;
    } else {
      k->outer_radius = (i + 0.5) * f->dr;
      if ((k->inner_radius = k->outer_radius - f->dr) < 0)
        k->inner_radius = 0;
      if (i == newCathodes - 1)
        k->outer_radius = d->outer_radius;
      if (k->outer_radius > d->outer_radius)
        k->outer_radius = d->outer_radius;
      if (k->inner_radius < d->inner_radius)
        k->inner_radius = d->inner_radius;
    }
    k->z_position = d->z_position;
    k->initial_vz = d->initial_pz / sqrt(1 + osqr(d->initial_pz)) * f->c_mks;
    k->initial_omega = d->initial_omega;
    k->current_density = d->current_density;
    k->temperature = temperature;
    k->work_function = d->work_function;
    k->field_emission = d->field_emission;
    k->field_emission_beta = d->field_emission_beta;
    k->start_time = d->start_time;
    k->stop_time = d->stop_time;
    k->time_offset = d->time_offset;
    k->autophase = d->autophase;
    k->number_per_step = d->number_per_step;
    k->QoMC = -d->e_mks / (d->me_mks * f->c_mks);
    k->correction_interval = d->distribution_correction_interval;
    k->start_step = -1;
    k->step_emission_started = -1;
    k->stiffness = d->stiffness;
    k->electrons_per_macroparticle = d->electrons_per_macroparticle;
    k->spread_over_dt = d->spread_over_dt;
    k->zoned_emission = d->zoned_emission;
    if ((k->discretize_radii = d->discretize_radii))
      k->correction_interval = 0;
    if (k->field_emission)
      k->zoned_emission = 0;
    if (!k->zoned_emission)
      k->correction_interval = 0;
    if ((k->halton_radix_dt = d->halton_radix_dt)) {
      if (!k->spread_over_dt)
        return bomb(err, errlen, "you have nonzero halton_radix_dt but didn't ask to spread_over_dt");
      if (d->distribution_correction_interval)
        return bomb(err, errlen, "you can't use distribution_correction_interval with halton sequence");
      if (d->halton_radix_dt < 2)
        return bomb(err, errlen, "halton_radix_dt<2");
    }
    if ((k->halton_radix_r = d->halton_radix_r)) {
      if (d->distribution_correction_interval)
        return bomb(err, errlen, "you can't use distribution_correction_interval with halton sequence");
      if (d->zoned_emission)
        return bomb(err, errlen, "you can't use zoned_emission with halton sequence");
      if (d->halton_radix_r < 2)
        return bomb(err, errlen, "halton_radix_r<2");
    }
    k->e_mks = d->e_mks;
    k->me_mks = d->me_mks;
    k->k_boltzmann_mks = d->k_boltzmann_mks;
    s->n++;
  }
  if (count)
    *count = (int32_t)newCathodes;
  return 0;
}

/* fieldEmissionCurrentDensity, src/cathode.c:799-810 */
double oracle_field_emission_current_density(double E, double beta, double wf) {
  double y, t, v;
  E = fabs(E) * beta;
  y = 3.79e-5 * sqrt(E) / wf;
  t = ((-0.0020 * y + 0.0131) * y + 0.1091) * y + 0.9899;
  v = ((0.0445 * y - 0.6782) * y - 0.4091) * y + 1.0437;
  return 1.54e-6 * osqr(E) / (wf * osqr(t)) * exp(-6.83e9 * v * pow(wf, 1.5) / E);
}

static double at(const double *a, const ofields *f, int64_t iz, int64_t ir) { return a[iz * (f->nr + 1) + ir]; }

/* set_current_density, src/cathode.c:673-715 (no on-axis field map) */
void oracle_set_current_density(ocathode_set *s, int32_t ic, const ofields *f) {
  ocathode *k = s->k + ic;
  int64_t iz = (int64_t)((k->z_position - f->zmin) / f->dz + 1.5);
  int64_t ir = (int64_t)((k->inner_radius + k->outer_radius) / 2 / f->dr);
  double Ez, wfCorrection;
  if (iz <= 0)
    iz = 1;
  if (!((k->temperature > 0 && k->work_function) || k->field_emission))
    return;
  Ez = f->constantEz;
  if (f->modes & OFL_TM_FIELDS)
    Ez += at(f->Ez, f, iz, ir) + at(f->EzImposed, f, iz, ir);
  if (Ez >= 0 && k->initial_vz <= 0)
    return;
  if (k->temperature > 0 && k->work_function) {
    wfCorrection = 0.012 * sqrt(fabs(Ez * 1e-5));
    if (wfCorrection > k->work_function)
      wfCorrection = k->work_function;
    k->current_density = 60e4 * osqr(k->temperature) *
                         exp(-(k->work_function - wfCorrection) / (k->temperature * k->k_boltzmann_mks / k->e_mks));
  } else if (k->field_emission) {
    k->current_density = oracle_field_emission_current_density(Ez, k->field_emission_beta, k->work_function);
  }
}

/* addThermalVelocities, src/cathode.c:761-772 */
static void add_thermal_velocities(orng *g, obeam *b, int64_t ip, double sigmaV) {
  double vx, vy, theta, sin_theta, cos_theta;
  vx = gauss_rn_lim(g, 0.0, sigmaV, 0.0);
  vy = gauss_rn_lim(g, 0.0, sigmaV, 0.0);
  theta = oracle_rng_uniform(g) * (2 * O_PI);
  sin_theta = sin(theta);
  cos_theta = cos(theta);
  b->vr[ip] += vx * cos_theta + vy * sin_theta;
  b->vphi[ip] += -vx * sin_theta + vy * cos_theta;
  b->vz[ip] += fabs(gauss_rn_lim(g, 0.0, sigmaV, 0.0));
}

/* the part of each emission loop body common to the three modes: state from
 * r, velocities and momenta, src/cathode.c:466-489 */
static void new_particle(ocathode_set *s, ocathode *k, obeam *b, const ofields *f, int64_t ip, double r) {
  double gamma;
  b->particleID[ip] = (int32_t)s->next_id++;
  b->rmin[ip] = b->zmin[ip] = DBL_MAX;
  b->rmax[ip] = b->zmax[ip] = -DBL_MAX;
  b->z[ip] = b->z0[ip] = k->z_position;
  b->r[ip] = b->r0[ip] = r;
  b->t0[ip] = f->time;
  b->generation[ip] = 0;
  b->Q[ip] = k->Q;
  b->vz[ip] = k->initial_vz;
  b->vr[ip] = 0;
  b->vphi[ip] = k->initial_omega * b->r[ip];
  if (k->sigmaVMaxwellian)
    add_thermal_velocities(&s->rng, b, ip, k->sigmaVMaxwellian);
  gamma = 1 / sqrt(1 - osqr(b->vz[ip] / f->c_mks) - osqr(b->vr[ip] / f->c_mks) - osqr(b->vphi[ip] / f->c_mks));
  b->pz[ip] = b->vz[ip] / f->c_mks * gamma;
  b->pr[ip] = b->vr[ip] / f->c_mks * gamma;
  b->r_pphi[ip] = b->r[ip] * b->vphi[ip] / f->c_mks * gamma;
  b->status[ip] = OPART_ACTIVE;
  b->npActive++;
}

/* drift and kick over dte with the fields at the emission point,
 * src/cathode.c:490-519 */
static void spread(ocathode *k, obeam *b, const ofields *f, int64_t ip, double dte) {
  int64_t iz, ir;
  double Ez, Er, pphi, gamma;
  iz = (int64_t)((k->z_position - f->zmin) / f->dz);
  if (iz <= 0)
    iz = 1;
  if (iz > f->nz)
    iz = f->nz;
  ir = (int64_t)(b->r[ip] / f->dr);
  if (ir > f->nr - 1) /* the library keeps the index inside the arrays */
    ir = f->nr - 1;
  Ez = f->constantEz;
  if (f->modes & OFL_TM_FIELDS)
    Ez = at(f->Ez, f, iz, ir) + at(f->EzImposed, f, iz, ir);
  ir += 1;
  iz -= 1;
  Er = f->constantEr;
  if (f->modes & OFL_TM_FIELDS)
    Er = at(f->Er, f, iz, ir) + at(f->ErImposed, f, iz, ir);
  b->t0[ip] -= dte;
  pphi = (b->r[ip] ? b->r_pphi[ip] / b->r[ip] : 0);
  b->z[ip] += b->vz[ip] * dte;
  b->z0[ip] = b->z[ip];
  b->r[ip] += b->vr[ip] * dte;
  b->r0[ip] = b->r[ip];
  b->pr[ip] += dte * k->QoMC * Er / k->stiffness;
  b->pz[ip] += dte * k->QoMC * Ez / k->stiffness;
  gamma = sqrt(osqr(b->pz[ip]) + osqr(b->pr[ip]) + osqr(pphi) + 1);
  b->vz[ip] = f->c_mks * b->pz[ip] / gamma;
  b->vr[ip] = f->c_mks * b->pr[ip] / gamma;
}

/* linear interpolation inside [t[0], t[n-1]] with a bisected bracket */
static double profile_at(const ocathode *k, double t) {
  int64_t lo = 0, hi = k->n_prof - 1;
  if (hi == 0)
    return k->A_prof[0];
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (t < k->t_prof[mid])
      hi = mid;
    else
      lo = mid;
  }
  return k->A_prof[lo] + (k->A_prof[hi] - k->A_prof[lo]) / (k->t_prof[hi] - k->t_prof[lo]) * (t - k->t_prof[lo]);
}

/* emit_electrons_from_cathode, src/cathode.c:239-671 */
int oracle_emit_electrons_from_cathode(ocathode_set *s, int32_t ic, obeam *b, const ofields *f, char *err,
                                       int64_t errlen) {
  ocathode *k = s->k + ic;
  int64_t ip, ib, np_new, np_start = b->np;
  double total, totalNominal, dr2, total_area, profile_factor, fnumber, area;
  k->last_count = 0;
  if (k->number_per_step == 0 && k->electrons_per_macroparticle == 0)
    return 0;
  if (k->start_step == -1) {
    if (k->current_density <= 1. / sqrt(DBL_MAX))
      return 0;
    k->ref_current_density = k->current_density;
    if (k->autophase) {
      k->stop_time = (k->stop_time - k->start_time) + f->time + k->time_offset;
      k->start_time = f->time + k->time_offset;
    }
    if (f->dti <= 0)
      return bomb(err, errlen, "integration step size is not defined (emit_electrons_from_cathode)");
    k->inner_r2 = osqr(k->inner_radius);
    k->outer_r2 = osqr(k->outer_radius);
    if ((area = O_PI * (k->outer_r2 - k->inner_r2)) != 0) {
      if (k->electrons_per_macroparticle) {
        k->number_per_step = area * f->dti * k->current_density / (k->e_mks * k->electrons_per_macroparticle);
        if (k->number_per_step == 0)
          return 0;
      } else
        k->electrons_per_macroparticle = area * f->dti * k->current_density / (k->e_mks * k->number_per_step);
    } else {
      k->correction_interval = 0;
      k->n_bins = 1;
    }
    k->start_step = (int64_t)((k->start_time - f->start_time) / f->dti);
    if (k->stop_time && k->stop_time > f->start_time)
      k->stop_step = (int64_t)((k->stop_time - f->start_time) / f->dti);
    else
      k->stop_step = LONG_MAX;
    k->Q = -k->e_mks * k->electrons_per_macroparticle;
    k->n_bins = (int64_t)((k->outer_radius - k->inner_radius) / f->dr + 1);
    free(k->distribution);
    free(k->deficit);
    free(k->area);
    free(k->rmin);
    k->distribution = (int64_t *)calloc((size_t)k->n_bins, sizeof(int64_t));
    k->deficit = (int64_t *)calloc((size_t)k->n_bins, sizeof(int64_t));
    k->area = (double *)calloc((size_t)k->n_bins, sizeof(double));
    k->rmin = (double *)calloc((size_t)k->n_bins, sizeof(double));
    k->n_emitted = 0;
    for (ib = 0, k->totalArea = 0; ib < k->n_bins; ib++) {
      int64_t ir = (int64_t)(k->inner_radius / f->dr + ib);
      if (ir == 0) {
        double r_max = f->dr / 2;
        k->area[ib] = osqr(r_max) * O_PI;
        k->rmin[ib] = 0;
      } else {
        double r_min = f->dr * (ir - 0.5), r_max = f->dr * (ir + 0.5);
        if ((k->area[ib] = O_PI * (osqr(r_max) - osqr(r_min))) < 0)
          return bomb(err, errlen, "negative area found for cathode annulus");
        k->rmin[ib] = r_min;
        k->totalArea += k->area[ib];
      }
    }
  }

  profile_factor = 1;
  if (k->n_prof) {
    if (f->time >= k->t_prof[0] && f->time <= k->t_prof[k->n_prof - 1])
      profile_factor = profile_at(k, f->time);
    else
      profile_factor = 0;
  }
  if (k->start_step > f->time_step || k->stop_step < f->time_step)
    return 0;
  if (k->step_emission_started < 0)
    k->step_emission_started = f->time_step;

  dr2 = (k->outer_r2 - k->inner_r2) / k->number_per_step;
  total_area = k->totalArea;
  total = totalNominal = k->number_per_step * (k->current_density / k->ref_current_density);
  np_new = b->np;
  if (k->correction_interval && f->time_step % k->correction_interval == 0) {
    total = 0;
    for (ib = 0; ib < k->n_bins; ib++) {
      if ((k->deficit[ib] = (int64_t)(((f->time_step - k->step_emission_started + 1) * (k->area[ib] / total_area) *
                                           totalNominal +
                                       0.5) -
                                      k->distribution[ib])) <= 0)
        k->deficit[ib] = 0;
      total += k->deficit[ib];
    }
  } else if (k->zoned_emission) {
    total = 0;
    for (ib = 0; ib < k->n_bins; ib++) {
      if (total_area) {
        fnumber = profile_factor * (k->area[ib] / total_area) * totalNominal;
        if (fnumber > 1) {
          k->deficit[ib] = (int64_t)fnumber;
          fnumber -= k->deficit[ib];
        } else
          k->deficit[ib] = 0;
        if (fnumber && fnumber > oracle_rng_uniform(&s->rng))
          k->deficit[ib] += 1;
        total += k->deficit[ib];
      }
    }
  } else {
    fnumber = total * profile_factor;
    total = (int64_t)fnumber;
    fnumber -= (int64_t)fnumber;
    if (fnumber && fnumber > oracle_rng_uniform(&s->rng))
      total += 1;
  }
  if ((np_new = (int64_t)(np_new + total)) < 0)
    return bomb(err, errlen, "negative number of particles created!");
  if (np_new > b->max_np)
    return bomb(err, errlen, "oracle beam capacity exceeded");

  if (total) {
    if (k->correction_interval) {
      for (ib = 0; ib < k->n_bins; ib++) {
        for (ip = b->np; ip < b->np + k->deficit[ib]; ip++) {
          new_particle(s, k, b, f, ip,
                       sqrt(oracle_rng_uniform(&s->rng) * k->area[ib] / O_PI + osqr(k->rmin[ib])));
          if (k->spread_over_dt)
            spread(k, b, f, ip, f->dti * (ip - b->np) / (1.0 * k->deficit[ib]));
        }
        k->distribution[ib] += k->deficit[ib];
        b->np = ip;
      }
    } else if (!k->discretize_radii) {
      for (ip = b->np; ip < b->np + total; ip++) {
        double rn;
        if (k->halton_radix_r)
          rn = oracle_radical_inverse(++k->halton_n_r, k->halton_radix_r);
        else
          rn = oracle_rng_uniform(&s->rng);
        new_particle(s, k, b, f, ip,
                     sqrt(rn * (osqr(k->outer_radius) - osqr(k->inner_radius)) + osqr(k->inner_radius)));
        if (k->spread_over_dt) {
          double dte;
          if (!k->halton_radix_dt)
            dte = f->dti * (ip - b->np) / (1.0 * total);
          else
            dte = f->dti * oracle_radical_inverse(++k->halton_n_dt, k->halton_radix_dt);
          spread(k, b, f, ip, dte);
        }
      }
      b->np = ip;
    } else {
      for (ip = b->np; ip < np_new; ip++) {
        new_particle(s, k, b, f, ip, sqrt(k->inner_r2 + (ip - b->np + 1) * dr2));
        if (k->spread_over_dt)
          spread(k, b, f, ip, f->dti * (ip - b->np) / (1.0 * (np_new - b->np)));
      }
    }
    k->last_count = np_new - np_start;
    k->n_emitted += k->last_count;
    b->np = np_new;
  }
  return 0;
}

int oracle_cathode_state(ocathode_set *s, int32_t ic, double *d8, int64_t *i5) {
  ocathode *k;
  if (!s || ic < 0 || ic >= s->n)
    return 1;
  k = s->k + ic;
  d8[0] = k->current_density;
  d8[1] = k->ref_current_density;
  d8[2] = k->number_per_step;
  d8[3] = k->electrons_per_macroparticle;
  d8[4] = k->Q;
  d8[5] = k->temperature;
  d8[6] = k->sigmaVMaxwellian;
  d8[7] = k->initial_vz;
  i5[0] = k->start_step;
  i5[1] = k->stop_step;
  i5[2] = k->n_bins;
  i5[3] = k->step_emission_started;
  i5[4] = s->n;
  return 0;
}
