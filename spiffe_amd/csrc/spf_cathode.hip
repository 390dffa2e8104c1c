// spf_cathode.hip -- cathode emission entry points.  The host part
// (spf_emit.cpp) decides how many macroparticles a cathode emits this step and
// makes their random draws in the reference's order; emit_kernel then writes
// each new particle at the end of the device arrays: momenta from the
// velocities, and the spread-over-dt drift and kick with the fields read where
// they live in HBM (src/cathode.c:458-656).
#include <cfloat>
#include <cmath>
#include <cstring>

#include "spf_internal.h"

using namespace spf;

namespace {

struct EmitArgs {
  int64_t n, cap;
  const double *in;   // 5 columns of n: r, dvr, dvphi, dvz, dte
  int thermal, spread, tm, iz, ldz, nr;
  double z_position, initial_vz, initial_omega, t0, Q, c, QoMC, stiffness, dr, constEz, constEr;
  int64_t first_id;
  const double *Ez, *Er, *EzImp, *ErImp;
  double *const *P;
  int32_t *pid;
  int16_t *gen;
  uint32_t *status;
  const int64_t *np;
};

// One new particle per thread, in emission order after the current np.
__global__ void __launch_bounds__(256) emit_kernel(EmitArgs a) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  int64_t dst = *a.np + k;
  if (dst >= a.cap) return;
  const double c = a.c;
  double r = a.in[k];
  double z = a.z_position, t0 = a.t0;
  double vz = a.initial_vz, vr = 0, vphi = a.initial_omega * r;
  if (a.thermal) {   // addThermalVelocities, src/cathode.c:769-771
    vr += a.in[a.n + k];
    vphi += a.in[2 * a.n + k];
    vz += a.in[3 * a.n + k];
  }
  double bz = vz / c, br = vr / c, bphi = vphi / c;
  double gamma = 1 / sqrt(1 - bz * bz - br * br - bphi * bphi);
  double pz = vz / c * gamma;
  double pr = vr / c * gamma;
  double rpphi = r * vphi / c * gamma;
  double z0 = z, r0 = r;
  if (a.spread) {
    // drift and kick over dte with the fields at the emission point,
    // src/cathode.c:494-519; the r index stays inside the arrays
    int64_t ir = (int64_t)(r / a.dr);
    if (ir > a.nr - 1) ir = a.nr - 1;
    if (ir < 0) ir = 0;
    double Ez = a.constEz, Er = a.constEr;
    if (a.tm) Ez = a.Ez[fix(a.ldz, a.iz, (int)ir)] + a.EzImp[fix(a.ldz, a.iz, (int)ir)];
    ir += 1;
    int iz1 = a.iz - 1;
    if (a.tm) Er = a.Er[fix(a.ldz, iz1, (int)ir)] + a.ErImp[fix(a.ldz, iz1, (int)ir)];
    double dte = a.in[4 * a.n + k];
    t0 -= dte;
    double pphi = r ? rpphi / r : 0;
    z += vz * dte;
    z0 = z;
    r += vr * dte;
    r0 = r;
    pr += dte * a.QoMC * Er / a.stiffness;
    pz += dte * a.QoMC * Ez / a.stiffness;
    gamma = sqrt(pz * pz + pr * pr + pphi * pphi + 1);
    vz = c * pz / gamma;
    vr = c * pr / gamma;
  }
  double *const *P = a.P;
  P[0][dst] = a.Q;
  P[1][dst] = z;
  P[2][dst] = r;
  P[3][dst] = pz;
  P[4][dst] = pr;
  P[5][dst] = rpphi;
  P[6][dst] = vz;
  P[7][dst] = vr;
  P[8][dst] = vphi;
  P[9][dst] = r0;
  P[10][dst] = z0;
  P[11][dst] = t0;
  P[12][dst] = -DBL_MAX;   // rmax
  P[13][dst] = DBL_MAX;    // rmin
  P[14][dst] = -DBL_MAX;   // zmax
  P[15][dst] = DBL_MAX;    // zmin
  a.pid[dst] = (int32_t)(a.first_id + k);
  a.gen[dst] = 0;
  a.status[dst] = SPF_PART_ACTIVE;
}

CathodeEnv env_of(const Run &R) {
  CathodeEnv e;
  e.zmin = R.d.zmin;
  e.zmax = R.d.zmax;
  e.rmax = R.d.rmax;
  e.dz = R.d.dz;
  e.dr = R.d.dr;
  e.dti = R.d.dt;
  e.start_time = R.d.start_time;
  e.time_step = R.time_step;
  e.time = R.time_step * R.d.dt + R.d.start_time;
  e.c_mks = R.d.c_mks;
  e.nz = R.d.nz;
  e.nr = R.d.nr;
  return e;
}

Cathode *get_cathode(Run *R, int32_t k) {
  if (!R || k < 0 || k >= (int32_t)R->cathodes.size()) return nullptr;
  return &R->cathodes[k];
}

}  // namespace

namespace spf {

// Emission of one cathode at the run's current step (also called by spf_step).
int cathode_emit_step(spf_ctx *c, int32_t run, int32_t cathode, int64_t *n_emitted) {
  Run *R = get_run(c, run);
  Cathode *K = get_cathode(R, cathode);
  if (!K) return fail(c, SPF_EINVAL, "bad run id or cathode index");
  if (n_emitted) *n_emitted = 0;
  EmitBatch b;
  std::string err;
  if (cathode_emit(*K, env_of(*R), R->emit_rng, b, err)) return fail(c, SPF_EINVAL, err);
  if (b.n == 0) return SPF_OK;
  if (R->np_ub + b.n > R->cap) {
    int rc = spf_particles_count(c, run, nullptr, nullptr);
    if (rc) return rc;
    if (R->np_ub + b.n > R->cap) return fail(c, SPF_ECAPACITY, "particle capacity exceeded by cathode emission");
  }
  size_t bytes = sizeof(double) * b.cols.size();
  HIPCALL(c, hipEventSynchronize(c->stage_done));
  if (c->hstage_bytes < bytes) {
    if (c->hstage) HIPCALL(c, hipHostFree(c->hstage));
    c->hstage = nullptr;
    HIPCALL(c, hipHostMalloc(&c->hstage, bytes, hipHostMallocDefault));
    c->hstage_bytes = bytes;
  }
  if (R->emit_cap < bytes) {
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (R->emit_stage) HIPCALL(c, hipFree(R->emit_stage));
    R->emit_stage = nullptr;
    HIPCALL(c, hipMalloc(&R->emit_stage, bytes));
    R->emit_cap = bytes;
  }
  memcpy(c->hstage, b.cols.data(), bytes);
  HIPCALL(c, hipMemcpyAsync(R->emit_stage, c->hstage, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipEventRecord(c->stage_done, c->stream));

  EmitArgs a{};
  a.n = b.n;
  a.cap = R->cap;
  a.in = R->emit_stage;
  a.thermal = b.thermal ? 1 : 0;
  a.spread = b.spread ? 1 : 0;
  a.tm = (R->d.modes & SPF_TM_FIELDS) ? 1 : 0;
  a.iz = b.iz_spread;
  a.ldz = R->ldz;
  a.nr = R->d.nr;
  a.z_position = K->z_position;
  a.initial_vz = K->initial_vz;
  a.initial_omega = K->initial_omega;
  a.t0 = R->time_step * R->d.dt + R->d.start_time;
  a.Q = K->Q;
  a.c = R->d.c_mks;
  a.QoMC = K->QoMC;
  a.stiffness = K->stiffness;
  a.dr = R->d.dr;
  a.constEz = R->d.constant_fields[0];
  a.constEr = R->d.constant_fields[1];
  a.first_id = R->next_pid;
  a.Ez = R->E2[0][R->cur[0]];
  a.Er = R->E2[1][R->cur[1]];
  a.EzImp = R->EzImp;
  a.ErImp = R->ErImp;
  a.P = R->Pdev;
  a.pid = R->pid;
  a.gen = R->gen;
  a.status = R->status;
  a.np = R->counters;
  hipLaunchKernelGGL(emit_kernel, dim3((unsigned)((b.n + 255) / 256)), dim3(256), 0, c->stream, a);
  HIPCALL(c, hipGetLastError());
  HIPCALL(c, launch_append_commit(*R, b.n, b.n, c->stream));
  R->np_ub += b.n;
  R->next_pid += b.n;
  if (n_emitted) *n_emitted = b.n;
  return SPF_OK;
}

int cathode_current_density(spf_ctx *c, int32_t run, int32_t cathode) {
  Run *R = get_run(c, run);
  Cathode *K = get_cathode(R, cathode);
  if (!K) return fail(c, SPF_EINVAL, "bad run id or cathode index");
  if (!cathode_needs_ez(*K)) return SPF_OK;   // no model reads Ez: nothing changes
  CathodeEnv env = env_of(*R);
  int64_t iz, ir;
  cathode_ez_index(*K, env, &iz, &ir);
  if (iz > R->d.nz) iz = R->d.nz;
  if (ir > R->d.nr) ir = R->d.nr;
  double Ez = R->d.constant_fields[0];
  if (R->d.modes & SPF_TM_FIELDS) {
    double *h = (double *)(c->pinned + 32);
    int64_t off = fix(R->ldz, (int)iz, (int)ir);
    HIPCALL(c, hipMemcpyAsync(h, R->E2[0][R->cur[0]] + off, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipMemcpyAsync(h + 1, R->EzImp + off, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    Ez += h[0] + h[1];
  }
  cathode_set_current_density(*K, Ez);
  return SPF_OK;
}

}  // namespace spf

extern "C" {

void spf_cathode_defaults(spf_cathode_desc *d) {
  if (!d) return;
  memset(d, 0, sizeof(*d));
  d->stiffness = 1;   // src/cathode.c:32-40
  d->zoned_emission = 1;
  d->random_number_seed = 987654321;
  d->field_emission_beta = 1;
}

int spf_cathode_define(spf_ctx *c, int32_t run, const spf_cathode_desc *d, int32_t *first, int32_t *count) {
  Run *R = get_run(c, run);
  if (!R || !d) return fail(c, SPF_EINVAL, "bad run id or null cathode description");
  if (d->e_mks > 0 && d->me_mks > 0) {
    double qomc = -d->e_mks / (d->me_mks * R->d.c_mks);
    if (qomc != R->d.QoMC) return fail(c, SPF_EINVAL, "cathode q/(m c) differs from the run's QoMC");
  }
  if (d->stiffness != R->d.stiffness) return fail(c, SPF_EINVAL, "cathode stiffness differs from the run's");
  size_t before = R->cathodes.size();
  std::string err;
  if (cathode_define(*d, env_of(*R), R->emit_rng, R->cathodes, err)) {
    R->cathodes.resize(before);
    return fail(c, SPF_EINVAL, err);
  }
  if (first) *first = (int32_t)before;
  if (count) *count = (int32_t)(R->cathodes.size() - before);
  return SPF_OK;
}

int spf_cathode_set_current_density(spf_ctx *c, int32_t run, int32_t cathode) {
  return cathode_current_density(c, run, cathode);
}

int spf_cathode_emit(spf_ctx *c, int32_t run, int32_t cathode, int64_t *n_emitted) {
  return cathode_emit_step(c, run, cathode, n_emitted);
}

int spf_cathode_state_get(spf_ctx *c, int32_t run, int32_t cathode, spf_cathode_state *out) {
  Run *R = get_run(c, run);
  Cathode *K = get_cathode(R, cathode);
  if (!K || !out) return fail(c, SPF_EINVAL, "bad run, cathode or output");
  out->current_density = K->current_density;
  out->ref_current_density = K->ref_current_density;
  out->number_per_step = K->number_per_step;
  out->electrons_per_macroparticle = K->electrons_per_macroparticle;
  out->Q = K->Q;
  out->temperature = K->temperature;
  out->sigma_v = K->sigmaV;
  out->inner_radius = K->inner_radius;
  out->outer_radius = K->outer_radius;
  out->initial_vz = K->initial_vz;
  out->start_step = K->start_step;
  out->stop_step = K->stop_step;
  out->n_bins = K->n_bins;
  out->n_emitted = K->n_emitted;
  out->last_count = K->last_count;
  return SPF_OK;
}

int spf_particle_id_next(spf_ctx *c, int32_t run, int64_t set_next, int64_t *next) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (set_next > 0) R->next_pid = set_next;
  if (next) *next = R->next_pid;
  return SPF_OK;
}

}  // extern "C"
