// spf_api_poisson.hip -- extern "C" entry points for the Poisson correction,
// the SOR solve, the imposed electrostatic field and the whole-step call.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "spf_internal.h"

using namespace spf;

namespace {

constexpr double PI_D = 3.141592653589793238462643;
constexpr int64_t kSorCheckBatches = 16;   // batches queued between host checks of the done flag
constexpr int kRing = 6;                   // Psi buffers the batches rotate through (> batches per check + 1)

// omega for half-sweep k: src/poisson.c:47-49 and :121-124
int ensure_omega(spf_ctx *c, Run *R, int64_t n) {
  if (R->omega && R->omega_n >= n) return SPF_OK;
  std::vector<double> w((size_t)n);
  double nz = R->d.nz, nr = R->d.nr, dz = R->d.dz, dr = R->d.dr;
  double r_jacobi = (cos(PI_D / nz) + (dz / dr) * (dz / dr) * cos(PI_D / nr)) / (1 + (dz / dr) * (dz / dr));
  double om = 1.0;
  for (int64_t k = 0; k < n; k++) {
    w[(size_t)k] = om;
    if (k == 0)
      om = 1.0 / (1 - r_jacobi * r_jacobi / 2);
    else
      om = 1.0 / (1 - r_jacobi * r_jacobi * om / 4);
  }
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (R->omega) HIPCALL(c, hipFree(R->omega));
  R->omega = nullptr;
  HIPCALL(c, hipMalloc((void **)&R->omega, sizeof(double) * (size_t)n));
  HIPCALL(c, hipMemcpyAsync(R->omega, w.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  R->omega_n = n;
  return SPF_OK;
}

// Makes `out` the solve's result buffer (*Psi) and the other ring buffers
// the run's spares.
void adopt_result(Run *R, double **Psi, double *const *ring, int nring, int out) {
  double *res = ring[out];
  double *spare[5];
  int si = 0;
  for (int i = 0; i < nring; i++)
    if (i != out) spare[si++] = ring[i];
  for (int i = 0; i < 5; i++) R->PsiSpare[i] = spare[i];
  *Psi = res;
}

// solve_poisson_cyl, src/poisson.c:17-130, on *Psi (a run-owned buffer that
// may be swapped with one of the run's spare buffers); *iters = n on
// convergence, -1 otherwise.  Batches rotate through a ring of kRing
// buffers: batch b reads ring[b % kRing] and writes ring[(b+1) % kRing].  The
// stopping rule is applied per check group of sor_check_group() batches by
// an extra workgroup of the next group's first batch launch (a check launch
// after the last group), on partial rows that alternate between two halves
// of the partials buffer.  After a stop the queued batches return at once;
// at most the rest of the stopping group and the next group's first batch
// have run past the stopping half-sweep, so the batch that contains it still
// has its input in the ring (kRing > batches per check + 1) and is re-run for
// exactly the half-sweeps needed.  Batches are queued in windows of
// kSorCheckBatches; the host reads a window's done flag while the next window
// is already queued, so the device never waits for the host between windows.
int sor_solve(spf_ctx *c, Run *R, double **Psi, const double *rhs, double accuracy, int64_t limit, int64_t *iters) {
  if (accuracy <= 0) return fail(c, SPF_EINVAL, "accuracy is <= 0 in solve_poisson_cyl");
  if (limit <= 1) return fail(c, SPF_EINVAL, "iteration_limit is <= 1 in solve_poisson_cyl");
  int rc = ensure_omega(c, R, limit);
  if (rc) return rc;
  hipStream_t s = c->stream;
  HIPCALL(c, launch_sor_reset(*R, s));
  HIPCALL(c, launch_sor_masks(*R, s));
  if (rhs)
    HIPCALL(c, launch_sum_abs(rhs, R->ldz, R->d.nz, R->d.nr - 1, R->red, R->red_n, R->sor_sums, s));
  const int K = sor_batch(), KC = sor_check_group();
  const int64_t group_rows = (int64_t)KC * K;   // partial rows of one check group
  double *ring[kRing] = {*Psi, R->PsiSpare[0], R->PsiSpare[1], R->PsiSpare[2], R->PsiSpare[3], R->PsiSpare[4]};
  int64_t nb = (limit + K - 1) / K, b = 0, converged = -1;
  int64_t *flag[2] = {c->pinned + 16, c->pinned + 18};   // [done, iterations] of each window in flight
  int pending = -1, win = 0;
  int64_t gi = 0;                                          // check groups queued so far
  int64_t prev_row = 0, prev_n0 = 0;                       // the queued group whose check is still due
  int prev_k = 0;
  auto collect = [&](int slot) -> int {
    HIPCALL(c, hipEventSynchronize(c->sor_ev[slot]));
    if (flag[slot][0]) converged = flag[slot][1];
    return SPF_OK;
  };
  while (b < nb && converged < 0) {
    int64_t win_end = std::min<int64_t>(nb, b + kSorCheckBatches);
    while (b < win_end) {
      const int64_t g0 = b, g1 = std::min<int64_t>(win_end, b + KC);
      const int64_t row0 = (gi & 1) * group_rows;
      {
        PhaseTimer bt(c, SPF_T_SOR_BATCH, g1 - b);   // the group's batch launches, back to back
        for (; b < g1; b++) {
          int64_t n0 = b * K;
          int k = (int)std::min<int64_t>(K, limit - n0);
          // the group's first batch also checks the previous group
          const int ck = b == g0 ? prev_k : 0;
          HIPCALL(c, launch_sor_batch(*R, ring[b % kRing], ring[(b + 1) % kRing], rhs, n0, k,
                                      (int)(row0 + (b - g0) * K), 0, s, prev_row, prev_n0, ck, accuracy));
          if (b == g0) prev_k = 0;
        }
      }
      prev_row = row0;
      prev_n0 = g0 * K;
      prev_k = (int)(std::min<int64_t>(g1 * K, limit) - g0 * K);
      gi++;
    }
    const int slot = win++ & 1;
    HIPCALL(c, hipMemcpyAsync(flag[slot], R->sor_state, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipEventRecord(c->sor_ev[slot], s));
    if (pending >= 0 && (rc = collect(pending))) return rc;
    pending = slot;
  }
  if (converged < 0) {
    // the last group's check, then the final flag
    if (prev_k > 0) HIPCALL(c, launch_sor_check(*R, prev_row, prev_n0, prev_k, accuracy, s));
    HIPCALL(c, hipMemcpyAsync(c->pinned + 20, R->sor_state, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    if (c->pinned[20]) converged = c->pinned[21];
  }
  int out;
  if (converged >= 0) {
    int64_t bc = converged / K, n0 = bc * K;
    int k = (int)(converged - n0 + 1);
    out = (int)((bc + 1) % kRing);
    // the batches after bc may have run before the check saw the stop: redo bc from its input
    HIPCALL(c, launch_sor_batch(*R, ring[bc % kRing], ring[out], rhs, n0, k, 0, 1, s));
    *iters = converged;
  } else {
    out = (int)(nb % kRing);
    *iters = -1;
  }
  adopt_result(R, Psi, ring, kRing, out);
  return SPF_OK;
}

}  // namespace

extern "C" {

int32_t spf_sor_batch_half_sweeps(void) { return spf::sor_batch(); }

int spf_poisson_setup(spf_ctx *c, int32_t run, const spf_poisson_params *p) {
  Run *R = get_run(c, run);
  if (!R || !p) return fail(c, SPF_EINVAL, "bad run or null parameters");
  // the checks of setup_poisson_correction, src/poisson_correction.c:51-70
  if (p->start_time < 0) return fail(c, SPF_EINVAL, "start_time < 0 for poisson_correction");
  if (p->step_interval < 0) return fail(c, SPF_EINVAL, "step_interval < 0 for poisson_correction");
  if (p->accuracy <= 0) return fail(c, SPF_EINVAL, "accuracy <= 0 for poisson_correction");
  if (p->maximum_iterations <= 0) return fail(c, SPF_EINVAL, "maximum iterations <= 0 for poisson_correction");
  if (p->guess_code < 0 || p->guess_code > 3) return fail(c, SPF_EINVAL, "unknown guess type for poisson_correction");
  if (p->guess_code == 1 || p->guess_code == 2)
    return fail(c, SPF_EUNSUPPORTED, "line-charge and point-charge guesses are not implemented on the device");
  if (p->error_charge_threshold < 0)
    return fail(c, SPF_EUNSUPPORTED, "relative error_charge_threshold is not implemented on the device");
  if (!(R->d.modes & SPF_TM_FIELDS)) return fail(c, SPF_EINVAL, "TM fields must be enabled in order to do poisson correction");
  R->pc = *p;
  R->pc_set = true;
  R->pc_start_step = -1;
  return SPF_OK;
}

int spf_poisson_solve(spf_ctx *c, int32_t run, int32_t which, double accuracy, int64_t max_iterations,
                      int64_t *iterations) {
  Run *R = get_run(c, run);
  if (!R || (which != 0 && which != 1)) return fail(c, SPF_EINVAL, "bad run or solve target");
  int64_t it = -1;
  int rc = which == 0 ? sor_solve(c, R, &R->Psi, R->Q, accuracy, max_iterations, &it)
                      : sor_solve(c, R, &R->PsiImp, nullptr, accuracy, max_iterations, &it);
  if (rc) return rc;
  if (iterations) *iterations = it;
  return SPF_OK;
}

int spf_poisson_sweeps(spf_ctx *c, int32_t run, int32_t which, int64_t n0, int64_t count) {
  Run *R = get_run(c, run);
  if (!R || (which != 0 && which != 1) || n0 < 0 || count < 0) return fail(c, SPF_EINVAL, "bad sweep arguments");
  if (count == 0) return SPF_OK;
  int rc = ensure_omega(c, R, n0 + count);
  if (rc) return rc;
  double **Psi = which == 0 ? &R->Psi : &R->PsiImp;
  const double *rhs = which == 0 ? R->Q : nullptr;
  const int K = sor_batch();
  HIPCALL(c, launch_sor_masks(*R, c->stream));
  double *buf[2] = {*Psi, R->PsiSpare[0]};
  int64_t nb = (count + K - 1) / K;
  {
    PhaseTimer bt(c, SPF_T_SOR_BATCH, nb);
    for (int64_t b = 0; b < nb; b++) {
      int k = (int)std::min<int64_t>(K, count - b * K);
      HIPCALL(c, launch_sor_batch(*R, buf[b % 2], buf[(b + 1) % 2], rhs, n0 + b * K, k, 0, 1, c->stream));
    }
  }
  if (nb % 2) {
    R->PsiSpare[0] = *Psi;
    *Psi = buf[1];
  }
  return SPF_OK;
}

int spf_poisson_correct(spf_ctx *c, int32_t run, int32_t space_charge, int64_t *iterations) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (iterations) *iterations = -2;
  const spf_poisson_params &p = R->pc;
  if (!R->pc_set || p.maximum_iterations <= 0) return SPF_OK;
  if (R->pc_start_step == -1) R->pc_start_step = (int64_t)((p.start_time - R->d.start_time) / R->d.dt);
  if (p.step_interval <= 0 || R->time_step % p.step_interval != 0 || R->pc_start_step > R->time_step) return SPF_OK;
  hipStream_t s = c->stream;
  PhaseTimer pt(c, SPF_T_POISSON);
  HIPCALL(c, launch_gauss_q(*R, s));
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s));
  if (p.error_charge_threshold > 0) {
    // src/poisson_correction.c:247-265 (Qerror summed in a fixed tree order)
    HIPCALL(c, launch_sum_plain(R->Q, R->ldz, R->d.nz, R->d.nr, R->red, R->red_n, R->sor_sums + 2, s));
    HIPCALL(c, hipMemcpyAsync(c->pinned + 8, R->sor_sums + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    double Qerror;
    memcpy(&Qerror, c->pinned + 8, sizeof(double));
    if (fabs(Qerror) < p.error_charge_threshold) return SPF_OK;
  }
  HIPCALL(c, launch_q_to_rhs(*R, s));
  if (p.guess_code == 3) HIPCALL(c, launch_zero_psi(*R, s));
  int64_t it = -1;
  int rc = sor_solve(c, R, &R->Psi, R->Q, p.accuracy, p.maximum_iterations, &it);
  if (rc) return rc;
  HIPCALL(c, launch_field_from_psi(*R, 0, 1.0, s));
  // recompute the apparent charge for output, :332-439
  HIPCALL(c, launch_gauss_q(*R, s));
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s));
  if (iterations) *iterations = it;
  return SPF_OK;
}

int spf_imposed_efield(spf_ctx *c, int32_t run, int32_t initialize, double ramp_time, double flat_top_time) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  R->ramp_time = ramp_time;
  R->flat_top_time = flat_top_time;
  const spf_poisson_params &p = R->pc;
  if (!R->pc_set || p.maximum_iterations <= 0) return SPF_OK;
  if (R->pc_start_step == -1) R->pc_start_step = (int64_t)((p.start_time - R->d.start_time) / R->d.dt);
  if (initialize) {
    int64_t it;
    int rc = sor_solve(c, R, &R->PsiImp, nullptr, p.accuracy, p.maximum_iterations, &it);
    if (rc) return rc;
  }
  // ramp, src/poisson_correction.c:504-517
  double ff = 1, T = ramp_time, F = flat_top_time, t = R->time;
  if (T) {
    if (t < T)
      ff = t / T;
    else if (t < (T + F))
      ff = 1;
    else if (t < (2 * T + F))
      ff = ((2 * T + F) - t) / T;
    else
      ff = 0;
    if (ff < 0) ff = 0;
  }
  if (ff == R->last_field_factor) return SPF_OK;
  R->last_field_factor = ff;
  R->imposed_e = true;
  HIPCALL(c, launch_field_from_psi(*R, 1, ff, c->stream));
  return SPF_OK;
}

int spf_step(spf_ctx *c, int32_t run, int32_t n_lines, const int32_t *line_dir, const int64_t *line_pos,
             const int64_t *line_start, const int64_t *line_end, const double *line_J) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (n_lines < 0 || (n_lines > 0 && (!line_dir || !line_pos || !line_start || !line_end || !line_J)))
    return fail(c, SPF_EINVAL, "bad line-current arguments");
  int rc;
  PhaseTimer pt(c, SPF_T_STEP);
  // src/integrate.c:136-138: the ramped imposed field uses the previous step's time
  if (R->ramp_time > 0 && (rc = spf_imposed_efield(c, run, 0, R->ramp_time, R->flat_top_time))) return rc;
  R->time = R->time_step * R->d.dt + R->d.start_time;
  for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)   // src/integrate.c:162-163
    if ((rc = cathode_emit_step(c, run, k, nullptr))) return rc;
  if (R->d.modes & SPF_SPACE_CHARGE) {
    // reset and deposition (src/integrate.c:169-173) in one pass over the
    // keys: the key kernel clears J, the folds add into it
    PhaseTimer dt(c, SPF_T_DEPOSIT);
    HIPCALL(c, deposit_currents(*R, c->stream, true));
  } else if ((rc = spf_currents_reset(c, run))) {
    return rc;
  }
  for (int32_t k = 0; k < n_lines; k++)
    if ((rc = spf_currents_add_line(c, run, line_dir[k], line_pos[k], line_start[k], line_end[k], line_J[k])))
      return rc;
  if (R->pc_set && (rc = spf_poisson_correct(c, run, (R->d.modes & SPF_SPACE_CHARGE) != 0, nullptr))) return rc;
  if ((rc = spf_fields_advance(c, run))) return rc;
  if ((rc = spf_particles_push(c, run, -1))) return rc;
  for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)   // src/integrate.c:190-191
    if ((rc = cathode_current_density(c, run, k))) return rc;
  R->time_step++;
  return SPF_OK;
}

}  // extern "C"
