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
//
// When a half-sweep's residual total lies within summation-order rounding of
// the threshold (sor_stop_rule), the device stops with the decision left to
// the host: the half-sweep is redone with the reference's expressions
// (launch_sor_exact_sweep), its |residual| and |rhs| are added serially in
// the reference's loop orders, and the comparison of src/poisson.c:125 is made
// on those sums.  If it does not stop there, the solve resumes at the next
// half-sweep from that Psi.
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
  int64_t start = 0;            // first half-sweep of this pass; ring[0] holds Psi before it
  bool drive_serial = false;    // sor_sums[9] holds the serial drive sum
  int out = 0;
  *iters = -1;
  for (;;) {
    int64_t nb = (limit - start + K - 1) / K, b = 0, converged = -1, ambiguous = 0;
    int64_t *flag[2] = {c->pinned + 40, c->pinned + 44};   // [done, iterations, host decides] per window
    int pending = -1, win = 0;
    int64_t gi = 0;                                          // check groups queued so far
    int64_t prev_row = 0, prev_n0 = 0;                       // the queued group whose check is still due
    int prev_k = 0;
    auto collect = [&](int slot) -> int {
      HIPCALL(c, hipEventSynchronize(c->sor_ev[slot]));
      if (flag[slot][0]) {
        converged = flag[slot][1];
        ambiguous = flag[slot][2];
      }
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
            int64_t n0 = start + b * K;
            int k = (int)std::min<int64_t>(K, limit - n0);
            // the group's first batch also checks the previous group
            const int ck = b == g0 ? prev_k : 0;
            HIPCALL(c, launch_sor_batch(*R, ring[b % kRing], ring[(b + 1) % kRing], rhs, n0, k,
                                        (int)(row0 + (b - g0) * K), 0, s, prev_row, prev_n0, ck, accuracy));
            if (b == g0) prev_k = 0;
          }
        }
        prev_row = row0;
        prev_n0 = start + g0 * K;
        prev_k = (int)(std::min<int64_t>(start + g1 * K, limit) - (start + g0 * K));
        gi++;
      }
      const int slot = win++ & 1;
      HIPCALL(c, hipMemcpyAsync(flag[slot], R->sor_state, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
      HIPCALL(c, hipEventRecord(c->sor_ev[slot], s));
      if (pending >= 0 && (rc = collect(pending))) return rc;
      pending = slot;
    }
    if (converged < 0) {
      // the last group's check, then the final flag
      if (prev_k > 0) HIPCALL(c, launch_sor_check(*R, prev_row, prev_n0, prev_k, accuracy, s));
      HIPCALL(c, hipMemcpyAsync(c->pinned + 48, R->sor_state, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
      HIPCALL(c, hipStreamSynchronize(s));
      if (c->pinned[48]) {
        converged = c->pinned[49];
        ambiguous = c->pinned[50];
      }
    }
    if (converged < 0) {
      out = (int)(nb % kRing);
      break;
    }
    const int64_t bc = (converged - start) / K, n0 = start + bc * K;
    out = (int)((bc + 1) % kRing);
    if (!ambiguous) {
      // the batches after bc may have run before the check saw the stop: redo bc from its input
      HIPCALL(c, launch_sor_batch(*R, ring[bc % kRing], ring[out], rhs, n0, (int)(converged - n0 + 1), 0, 1, s));
      *iters = converged;
      break;
    }
    // Psi before half-sweep `converged` into ring[out], then that half-sweep
    // with its residuals dumped, and the serial sums of the reference
    if (converged > n0)
      HIPCALL(c, launch_sor_batch(*R, ring[bc % kRing], ring[out], rhs, n0, (int)(converged - n0), 0, 1, s));
    else
      HIPCALL(c, hipMemcpyAsync(ring[out], ring[bc % kRing], sizeof(double) * (size_t)R->cells,
                                hipMemcpyDeviceToDevice, s));
    HIPCALL(c, launch_sor_exact_sweep(*R, ring[out], rhs, converged, R->stage, s));
    HIPCALL(c, launch_serial_sum_ir_iz(R->stage, R->d.nz, R->d.nz, R->d.nr - 1, R->sor_sums + 8, s));
    if (rhs && !drive_serial)
      HIPCALL(c, launch_serial_sum_iz_ir(rhs, R->ldz, R->d.nz, R->d.nr - 1, 1, R->sor_sums + 9, s));
    drive_serial = rhs != nullptr;
    HIPCALL(c, hipMemcpyAsync(c->pinned + 52, R->sor_sums + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    double res, drive = 0;
    memcpy(&res, c->pinned + 52, sizeof(double));
    if (rhs) memcpy(&drive, c->pinned + 53, sizeof(double));
    if (res < accuracy * drive || res == 0) {   // src/poisson.c:125-126, converged > 0
      *iters = converged;
      break;
    }
    if (converged + 1 >= limit) break;   // that was the last half-sweep: not converged
    // resume at the next half-sweep from ring[out]: it becomes ring[0]
    double *rot[kRing];
    for (int i = 0; i < kRing; i++) rot[i] = ring[(out + i) % kRing];
    for (int i = 0; i < kRing; i++) ring[i] = rot[i];
    out = 0;
    start = converged + 1;
    HIPCALL(c, launch_sor_state_clear(*R, s));
  }
  adopt_result(R, Psi, ring, kRing, out);
  return SPF_OK;
}

// The error-charge test of src/poisson_correction.c:247-265:
// skip = (thr > 0 && |Qerror| < thr) || (thr < 0 && (Qreal == 0 || |Qerror/Qreal| < |thr|)),
// with Qerror the serial sum of Q over iz < nz, ir < nr (iz outer) and Qreal the
// serial sum of the removed charges in particle order.  The device first
// decides from fixed-order tree sums and their sums of magnitudes, which bound
// how far any other summation order can land (order_margin); only when the
// test could go either way are the serial sums formed on the device.
int error_charge_skips(spf_ctx *c, Run *R, bool space_charge, double thr, bool *skip) {
  hipStream_t s = c->stream;
  HIPCALL(c, launch_sum2_region(R->Q, R->ldz, R->d.nz, R->d.nr, R->red, R->red_n, R->sor_sums + 2, s));
  HIPCALL(c, hipMemcpyAsync(c->pinned + 52, R->sor_sums + 2, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCALL(c, hipStreamSynchronize(s));
  double h[4];
  memcpy(h, c->pinned + 52, sizeof(h));
  const double Te = h[0], Ae = h[1];
  const double Tr = space_charge ? -h[2] : 0.0, Ar = space_charge ? h[3] : 0.0;
  const double de = order_margin((int64_t)R->d.nz * R->d.nr, Ae), dr = order_margin(R->np_ub, Ar);
  const double big = 0x1p1000, slack = 1e-12;
  // 1: skip, 0: correct, -1: undecided
  int dec = -1;
  const bool finite = std::isfinite(Te) && std::isfinite(Ae) && Ae < big && std::isfinite(Tr) && Ar < big &&
                      std::fabs(thr) < big && std::fabs(thr) > 0x1p-1000;
  if (thr < 0 && !space_charge) {
    dec = 1;   // Qreal is 0
  } else if (finite) {
    const double elo = std::max(std::fabs(Te) - de, 0.0), ehi = std::fabs(Te) + de;
    if (thr > 0) {
      if (ehi * (1 + slack) < thr)
        dec = 1;
      else if (elo * (1 - slack) >= thr)
        dec = 0;
    } else if (Ar == 0) {
      dec = 1;   // every removed charge is +-0, so Qreal is +0 in any order
    } else if (std::fabs(Tr) > dr * (1 + slack)) {
      const double rlo = std::fabs(Tr) - dr, rhi = std::fabs(Tr) + dr;
      const double qhi = ehi / rlo * (1 + slack), qlo = elo / rhi * (1 - slack);
      if (qhi < std::fabs(thr))
        dec = 1;
      else if (qlo >= std::fabs(thr))
        dec = 0;
    }
  }
  if (dec < 0) {
    HIPCALL(c, launch_serial_sum_iz_ir(R->Q, R->ldz, R->d.nz, R->d.nr, 0, R->sor_sums + 6, s));
    if (thr < 0) HIPCALL(c, launch_qreal_serial(*R, R->sor_sums + 7, s));
    HIPCALL(c, hipMemcpyAsync(c->pinned + 52, R->sor_sums + 6, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    double Qerror, Qreal = 0;
    memcpy(&Qerror, c->pinned + 52, sizeof(double));
    if (thr < 0) memcpy(&Qreal, c->pinned + 53, sizeof(double));
    dec = (thr > 0 && fabs(Qerror) < thr) || (thr < 0 && (Qreal == 0 || fabs(Qerror / Qreal) < fabs(thr)));
  }
  *skip = dec == 1;
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
  const double thr = p.error_charge_threshold;
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s, thr < 0 ? R->sor_sums + 4 : nullptr));
  if (thr != 0) {
    bool skip = false;
    int rc = error_charge_skips(c, R, space_charge != 0, thr, &skip);
    if (rc) return rc;
    if (skip) return SPF_OK;
  }
  HIPCALL(c, launch_q_to_rhs(*R, s));
  if (p.guess_code == 3) HIPCALL(c, launch_zero_psi(*R, s));
  int64_t it = -1;
  int rc = sor_solve(c, R, &R->Psi, R->Q, p.accuracy, p.maximum_iterations, &it);
  if (rc) return rc;
  R->pc_last_iters = it;
  R->pc_solves++;
  R->pc_half_sweeps += it >= 0 ? it + 1 : p.maximum_iterations;
  HIPCALL(c, launch_field_from_psi(*R, 0, 1.0, s));
  // recompute the apparent charge for output, :332-439
  HIPCALL(c, launch_gauss_q(*R, s));
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s));
  if (iterations) *iterations = it;
  return SPF_OK;
}

int spf_poisson_stats(spf_ctx *c, int32_t run, int64_t *last_iterations, int64_t *solves, int64_t *half_sweeps) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (last_iterations) *last_iterations = R->pc_last_iters;
  if (solves) *solves = R->pc_solves;
  if (half_sweeps) *half_sweeps = R->pc_half_sweeps;
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
  if ((rc = spf_particles_push(c, run, R->z_forces_start))) return rc;   // src/integrate.c:186-188
  for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)   // src/integrate.c:190-191
    if ((rc = cathode_current_density(c, run, k))) return rc;
  R->time_step++;
  return SPF_OK;
}

}  // extern "C"
