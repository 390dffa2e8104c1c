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

// One run's solve of solve_poisson_cyl, src/poisson.c:17-130.  Batches
// rotate through a ring of kRing buffers: batch b of a pass reads
// ring[b % kRing] and writes ring[(b+1) % kRing], ring[0] holding Psi before
// the pass's first half-sweep `start`.  The stopping rule is applied per check
// group of sor_check_group() batches by an extra workgroup of the next
// group's first batch launch (a check launch after the last group), on
// partial rows that alternate between two halves of the partials buffer.
// After a stop the queued batches return at once; at most the rest of the
// stopping group and the next group's first batch have run past the stopping
// half-sweep, so the batch that contains it still has its input in the ring
// (kRing > batches per check + 1) and is re-run for exactly the half-sweeps
// needed.  Batches are queued in windows of kSorCheckBatches; the host reads a
// window's done flags while the next window is already queued, so the device
// never waits for the host between windows.
//
// When a half-sweep's residual total lies within summation-order rounding of
// the threshold (sor_stop_rule), the device stops with the decision left to
// the host: the half-sweep is redone with the reference's expressions
// (launch_sor_exact_sweep), its |residual| and |rhs| are added serially in
// the reference's loop orders, and the comparison of src/poisson.c:125 is made
// on those sums.  If it does not stop there, the solve resumes at the next
// half-sweep from that Psi.
struct SorPass {
  Run *R = nullptr;
  const double *rhs = nullptr;
  double accuracy = 0;
  int64_t limit = 0;
  double *ring[kRing] = {};
  int64_t start = 0;
  bool drive_serial = false;   // sor_sums[9] holds the serial drive sum
  int out = 0;                 // ring index of the result once done
  int64_t iters = -1;
  bool done = false;
  int64_t converged = -1, ambiguous = 0, nb = 0;   // what the last pass of batches reported
};

int sor_pass_init(spf_ctx *c, SorPass &p, Run *R, double *Psi, const double *rhs, double accuracy, int64_t limit,
                  int shape) {
  if (accuracy <= 0) return fail(c, SPF_EINVAL, "accuracy is <= 0 in solve_poisson_cyl");
  if (limit <= 1) return fail(c, SPF_EINVAL, "iteration_limit is <= 1 in solve_poisson_cyl");
  int rc = ensure_omega(c, R, limit);
  if (rc) return rc;
  R->sor_shape = shape;   // the masks below and every launch of this solve use it
  hipStream_t s = c->stream;
  HIPCALL(c, launch_sor_reset(*R, s));
  HIPCALL(c, launch_sor_masks(*R, s));
  if (rhs) HIPCALL(c, launch_sum_abs(rhs, R->ldz, R->d.nz, R->d.nr - 1, R->red, R->red_n, R->sor_sums, s));
  p = SorPass();
  p.R = R;
  p.rhs = rhs;
  p.accuracy = accuracy;
  p.limit = limit;
  p.ring[0] = Psi;
  for (int i = 1; i < kRing; i++) p.ring[i] = R->PsiSpare[i - 1];
  return SPF_OK;
}

// Queues the batches of half-sweeps [start, limit) of the passes in `ps`
// (which share start and limit) until each has stopped: one run by its own
// launches, several by ragged launches that advance them in lockstep.  Sets
// each pass's converged (-1: ran to the limit), ambiguous and nb.
int sor_run_batches(spf_ctx *c, SorPass *const *ps, int np) {
  hipStream_t s = c->stream;
  const int K = sor_batch(*ps[0]->R), KC = sor_check_group(*ps[0]->R);
  const int64_t group_rows = (int64_t)KC * K;   // partial rows of one check group
  const int64_t start = ps[0]->start, limit = ps[0]->limit;
  const int64_t nb = (limit - start + K - 1) / K;
  const double accuracy = ps[0]->accuracy;
  // ragged launches: the per-run table once, and per-window flags of each run
  const void *tdev = nullptr;
  const int32_t *first_dev = nullptr;
  int64_t total_blocks = 0;
  if (np > 1) {
    std::vector<Run *> runs(np);
    std::vector<const double *> rhs(np);
    double *rings[64][6];
    if (np > 64) return fail(c, SPF_EINVAL, "at most 64 runs solve in one ragged batch");
    for (int i = 0; i < np; i++) {
      runs[i] = ps[i]->R;
      rhs[i] = ps[i]->rhs;
      for (int j = 0; j < 6; j++) rings[i][j] = ps[i]->ring[j];
    }
    void *host = nullptr, *dev = nullptr;
    int rc = batch_table(c, sor_multi_table_bytes(np), &host, &dev);
    if (rc) return rc;
    total_blocks = sor_multi_table(runs.data(), rhs.data(), rings, np, accuracy, host, dev, &first_dev);
    HIPCALL(c, batch_table_upload(c, host, dev, sor_multi_table_bytes(np)));
    tdev = dev;
  }
  int rc = batch_flags(c, (size_t)np * 3 * 2);
  if (rc) return rc;
  int64_t *flags = c->bw_flags;   // [slot][run][done, iterations, host decides]
  std::vector<char> live(np, 1);
  int n_live = np;
  for (int i = 0; i < np; i++) {
    ps[i]->converged = -1;
    ps[i]->ambiguous = 0;
    ps[i]->nb = nb;
  }
  int64_t b = 0;
  int pending = -1, win = 0;
  int64_t gi = 0;                       // check groups queued so far
  int64_t prev_row = 0, prev_n0 = 0;    // the queued group whose check is still due
  int prev_k = 0;
  auto collect = [&](int slot) -> int {
    HIPCALL(c, hipEventSynchronize(c->sor_ev[slot]));
    for (int i = 0; i < np; i++) {
      const int64_t *f = flags + ((int64_t)slot * np + i) * 3;
      if (live[i] && f[0]) {
        ps[i]->converged = f[1];
        ps[i]->ambiguous = f[2];
        live[i] = 0;
        n_live--;
      }
    }
    return SPF_OK;
  };
  while (b < nb && n_live > 0) {
    int64_t win_end = std::min<int64_t>(nb, b + kSorCheckBatches);
    while (b < win_end) {
      const int64_t g0 = b, g1 = std::min<int64_t>(win_end, b + KC);
      const int64_t row0 = (gi & 1) * group_rows;
      {
        PhaseTimer bt(c, SPF_T_SOR_BATCH, g1 - b);   // the group's batch launches, back to back
        for (; b < g1; b++) {
          int64_t n0 = start + b * K;
          int k = (int)std::min<int64_t>(K, limit - n0);
          const int ck = b == g0 ? prev_k : 0;   // the group's first batch also checks the previous group
          const int64_t prow = row0 + (b - g0) * K;
          if (np == 1) {
            HIPCALL(c, launch_sor_batch(*ps[0]->R, ps[0]->ring[b % kRing], ps[0]->ring[(b + 1) % kRing], ps[0]->rhs,
                                        n0, k, (int)prow, 0, s, prev_row, prev_n0, ck, accuracy));
          } else {
            HIPCALL(c, launch_sor_multi(ps[0]->R->sor_shape, tdev, first_dev, np, total_blocks, n0, b, k, prow, prev_row,
                                        prev_n0, ck, s));
          }
          if (b == g0) prev_k = 0;
        }
      }
      prev_row = row0;
      prev_n0 = start + g0 * K;
      prev_k = (int)(std::min<int64_t>(start + g1 * K, limit) - (start + g0 * K));
      gi++;
    }
    const int slot = win++ & 1;
    for (int i = 0; i < np; i++)
      if (live[i])
        HIPCALL(c, hipMemcpyAsync(flags + ((int64_t)slot * np + i) * 3, ps[i]->R->sor_state, 3 * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipEventRecord(c->sor_ev[slot], s));
    if (pending >= 0 && (rc = collect(pending))) return rc;
    pending = slot;
  }
  if (pending >= 0 && (rc = collect(pending))) return rc;
  if (n_live > 0) {
    // the last group's check of the runs still going, then their final flags
    for (int i = 0; i < np; i++)
      if (live[i]) {
        if (prev_k > 0) HIPCALL(c, launch_sor_check(*ps[i]->R, prev_row, prev_n0, prev_k, accuracy, s));
        HIPCALL(c, hipMemcpyAsync(flags + (int64_t)i * 3, ps[i]->R->sor_state, 3 * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
      }
    HIPCALL(c, hipStreamSynchronize(s));
    for (int i = 0; i < np; i++)
      if (live[i] && flags[(int64_t)i * 3]) {
        ps[i]->converged = flags[(int64_t)i * 3 + 1];
        ps[i]->ambiguous = flags[(int64_t)i * 3 + 2];
      }
  }
  return SPF_OK;
}

// Turns what the batches reported into the result (p.done, p.iters, p.out),
// or, after a half-sweep the host decided was not the last, rotates the ring
// so that the solve resumes at the next half-sweep.
int sor_settle(spf_ctx *c, SorPass &p) {
  hipStream_t s = c->stream;
  Run *R = p.R;
  const int K = sor_batch(*R);
  const int64_t converged = p.converged;
  if (converged < 0) {
    p.out = (int)(p.nb % kRing);
    p.done = true;
    return SPF_OK;
  }
  const int64_t bc = (converged - p.start) / K, n0 = p.start + bc * K;
  const int out = (int)((bc + 1) % kRing);
  p.out = out;
  if (!p.ambiguous) {
    // the batches after bc may have run before the check saw the stop: redo bc from its input
    HIPCALL(c, launch_sor_batch(*R, p.ring[bc % kRing], p.ring[out], p.rhs, n0, (int)(converged - n0 + 1), 0, 1, s));
    p.iters = converged;
    p.done = true;
    return SPF_OK;
  }
  // Psi before half-sweep `converged` into ring[out], then that half-sweep
  // with its residuals dumped, and the serial sums of the reference
  if (converged > n0)
    HIPCALL(c, launch_sor_batch(*R, p.ring[bc % kRing], p.ring[out], p.rhs, n0, (int)(converged - n0), 0, 1, s));
  else
    HIPCALL(c, hipMemcpyAsync(p.ring[out], p.ring[bc % kRing], sizeof(double) * (size_t)R->cells,
                              hipMemcpyDeviceToDevice, s));
  HIPCALL(c, launch_sor_exact_sweep(*R, p.ring[out], p.rhs, converged, R->stage, s));
  HIPCALL(c, launch_serial_sum_ir_iz(R->stage, R->d.nz, R->d.nz, R->d.nr - 1, R->sor_sums + 8, s));
  if (p.rhs && !p.drive_serial)
    HIPCALL(c, launch_serial_sum_iz_ir(p.rhs, R->ldz, R->d.nz, R->d.nr - 1, 1, R->sor_sums + 9, s));
  p.drive_serial = p.rhs != nullptr;
  HIPCALL(c, hipMemcpyAsync(c->pinned + 52, R->sor_sums + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCALL(c, hipStreamSynchronize(s));
  double res, drive = 0;
  memcpy(&res, c->pinned + 52, sizeof(double));
  if (p.rhs) memcpy(&drive, c->pinned + 53, sizeof(double));
  if (res < p.accuracy * drive || res == 0) {   // src/poisson.c:125-126, converged > 0
    p.iters = converged;
    p.done = true;
    return SPF_OK;
  }
  if (converged + 1 >= p.limit) {   // that was the last half-sweep: not converged
    p.done = true;
    return SPF_OK;
  }
  // resume at the next half-sweep from ring[out]: it becomes ring[0]
  double *rot[kRing];
  for (int i = 0; i < kRing; i++) rot[i] = p.ring[(out + i) % kRing];
  for (int i = 0; i < kRing; i++) p.ring[i] = rot[i];
  p.out = 0;
  p.start = converged + 1;
  HIPCALL(c, launch_sor_state_clear(*R, s));
  return SPF_OK;
}

// runs a pass to its end on its own (the single-run solve, and any run of a
// ragged solve that resumes after a host decision)
int sor_finish(spf_ctx *c, SorPass &p) {
  int rc;
  while (!p.done) {
    SorPass *one = &p;
    if ((rc = sor_run_batches(c, &one, 1)) || (rc = sor_settle(c, p))) return rc;
  }
  return SPF_OK;
}

int sor_solve(spf_ctx *c, Run *R, double **Psi, const double *rhs, double accuracy, int64_t limit, int64_t *iters) {
  SorPass p;
  const int shape = c->sor_shape >= 0 ? c->sor_shape : sor_shape_for((int64_t)(R->d.nz + 1) * (R->d.nr + 1));
  int rc = sor_pass_init(c, p, R, *Psi, rhs, accuracy, limit, shape);
  if (rc || (rc = sor_finish(c, p))) return rc;
  *iters = p.iters;
  adopt_result(R, Psi, p.ring, kRing, p.out);
  return SPF_OK;
}

// The solves of several runs at once (same accuracy and limit): ragged
// launches advance them in lockstep until each stops; a run whose stop the
// host decides against resumes on its own.  Each run's half-sweeps, partial
// sums and stopping decisions are those of its single-run solve.
int sor_solve_multi(spf_ctx *c, Run *const *runs, const double *const *rhs, double accuracy, int64_t limit, int np,
                    int64_t *iters) {
  std::vector<SorPass> ps(np);
  std::vector<SorPass *> pp(np);
  int rc;
  // one tile shape for the whole ragged solve, chosen by the points it covers:
  // together the runs fill the GPU as one large grid would
  int64_t points = 0;
  for (int i = 0; i < np; i++) points += (int64_t)(runs[i]->d.nz + 1) * (runs[i]->d.nr + 1);
  const int shape = c->sor_shape >= 0 ? c->sor_shape : sor_shape_for(points);
  for (int i = 0; i < np; i++) {
    if ((rc = sor_pass_init(c, ps[i], runs[i], runs[i]->Psi, rhs[i], accuracy, limit, shape))) return rc;
    pp[i] = &ps[i];
  }
  for (int i0 = 0; i0 < np; i0 += 64) {
    const int n = std::min(64, np - i0);
    if ((rc = sor_run_batches(c, pp.data() + i0, n))) return rc;
  }
  for (int i = 0; i < np; i++) {
    if ((rc = sor_settle(c, ps[i])) || (rc = sor_finish(c, ps[i]))) return rc;
    iters[i] = ps[i].iters;
    adopt_result(runs[i], &runs[i]->Psi, ps[i].ring, kRing, ps[i].out);
  }
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
  const double de = order_margin((int64_t)R->d.nz * R->d.nr, Ae),
               dr = order_margin(R->np_ub * deposit_slots_per_particle(R->d.modes), Ar);
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

// Whether add_poisson_correction (src/poisson_correction.c:73-93) corrects at the run's current step.
bool poisson_due(Run *R) {
  const spf_poisson_params &p = R->pc;
  if (!R->pc_set || p.maximum_iterations <= 0) return false;
  if (R->pc_start_step == -1) R->pc_start_step = (int64_t)((p.start_time - R->d.start_time) / R->d.dt);
  return !(p.step_interval <= 0 || R->time_step % p.step_interval != 0 || R->pc_start_step > R->time_step);
}

// The correction up to its solve (src/poisson_correction.c:141-287): the
// Gauss-law charge, the particles' charge taken off, the error-charge test,
// the right-hand side and the guess.  *solve: the solve of Psi is due.
int poisson_pre(spf_ctx *c, Run *R, int space_charge, bool *solve) {
  *solve = false;
  if (!poisson_due(R)) return SPF_OK;
  const spf_poisson_params &p = R->pc;
  hipStream_t s = c->stream;
  const int nz = R->d.nz, nr = R->d.nr;
  HIPCALL(c, launch_gauss_q(*R, s));
  // the point-charge guess's "centroid": the sum of the Gauss-law charges
  // times their z, ir outer and iz inner as it is formed at :176 (metal and
  // Neumann points hold +0 and add nothing)
  if (p.guess_code == 2)
    HIPCALL(c, launch_serial_moment_ir_iz(R->Q, R->ldz, nz, nr, R->d.dz, R->d.zmin, R->sor_sums + 13, s));
  const double thr = p.error_charge_threshold;
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s, thr < 0 ? R->sor_sums + 4 : nullptr));
  if (thr != 0) {
    bool skip = false;
    int rc = error_charge_skips(c, R, space_charge != 0, thr, &skip);
    if (rc) return rc;
    if (skip) return SPF_OK;
  }
  // the guesses take Qerror, the serial sum of :246-250 (iz outer)
  if (p.guess_code == 1 || p.guess_code == 2)
    HIPCALL(c, launch_serial_sum_iz_ir(R->Q, R->ldz, nz, nr, 0, R->sor_sums + 12, s));
  HIPCALL(c, launch_q_to_rhs(*R, s));
  if (p.guess_code == 3) {
    HIPCALL(c, launch_zero_psi(*R, s));
  } else if (p.guess_code == 2) {
    HIPCALL(c, launch_point_guess(*R, R->sor_sums + 12, R->sor_sums + 13, s));
  } else if (p.guess_code == 1) {
    // line charge, :575-582: Qerror/(2 pi eps0) * log(r) per row, with the
    // host's log (the reference's libm), r = 0.5*dr on the axis row
    HIPCALL(c, hipMemcpyAsync(c->pinned + 54, R->sor_sums + 12, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    double charge;
    memcpy(&charge, c->pinned + 54, sizeof(double));
    std::vector<double> row(nr);
    for (int ir = 0; ir < nr; ir++)
      row[ir] = charge / (2 * 3.141592653589793238462643 * R->d.epsilon_o) * std::log((ir ? (double)ir : 0.5) * R->d.dr);
    HIPCALL(c, hipMemcpyAsync(R->stage, row.data(), nr * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCALL(c, launch_line_guess(*R, R->stage, s));
    HIPCALL(c, hipStreamSynchronize(s));
  }
  *solve = true;
  return SPF_OK;
}

// After the solve (src/poisson_correction.c:288-439): E corrected by the
// gradient of Psi, the apparent charge recomputed for output.
int poisson_post(spf_ctx *c, Run *R, int space_charge, int64_t it) {
  hipStream_t s = c->stream;
  R->pc_last_iters = it;
  R->pc_solves++;
  R->pc_half_sweeps += it >= 0 ? it + 1 : R->pc.maximum_iterations;
  HIPCALL(c, launch_field_from_psi(*R, 0, 1.0, s));
  HIPCALL(c, launch_gauss_q(*R, s));
  if (space_charge) HIPCALL(c, launch_q_subtract(*R, s));
  return SPF_OK;
}

}  // namespace

extern "C" {

int spf_sor_batch_half_sweeps(spf_ctx *c, int32_t run, int32_t *half_sweeps) {
  Run *R = get_run(c, run);
  if (!R || !half_sweeps) return fail(c, SPF_EINVAL, "bad run id or null argument");
  *half_sweeps = sor_batch(*R);
  return SPF_OK;
}

int spf_poisson_setup(spf_ctx *c, int32_t run, const spf_poisson_params *p) {
  Run *R = get_run(c, run);
  if (!R || !p) return fail(c, SPF_EINVAL, "bad run or null parameters");
  // the checks of setup_poisson_correction, src/poisson_correction.c:51-70
  if (p->start_time < 0) return fail(c, SPF_EINVAL, "start_time < 0 for poisson_correction");
  if (p->step_interval < 0) return fail(c, SPF_EINVAL, "step_interval < 0 for poisson_correction");
  if (p->accuracy <= 0) return fail(c, SPF_EINVAL, "accuracy <= 0 for poisson_correction");
  if (p->maximum_iterations <= 0) return fail(c, SPF_EINVAL, "maximum iterations <= 0 for poisson_correction");
  if (p->guess_code < 0 || p->guess_code > 3) return fail(c, SPF_EINVAL, "unknown guess type for poisson_correction");
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
  const int K = sor_batch(*R);
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
  bool solve = false;
  PhaseTimer pt(c, SPF_T_POISSON, poisson_due(R) ? 1 : 0);
  int rc = poisson_pre(c, R, space_charge, &solve);
  if (rc || !solve) return rc;
  int64_t it = -1;
  if ((rc = sor_solve(c, R, &R->Psi, R->Q, R->pc.accuracy, R->pc.maximum_iterations, &it))) return rc;
  if ((rc = poisson_post(c, R, space_charge, it))) return rc;
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
  if (R->cell_sort > 0 && R->time_step % R->cell_sort == 0 && (rc = spf_particles_sort(c, run))) return rc;
  // with the J time filter, J is scaled instead of cleared (src/integrate.c:168-171)
  const bool filt = R->jfm != 0;
  if (filt && (rc = spf_currents_scale(c, run, R->jm1))) return rc;
  if (R->d.modes & SPF_SPACE_CHARGE) {
    // reset and deposition (src/integrate.c:169-173) in one pass over the
    // keys: the key kernel clears J, the folds add into it
    PhaseTimer dt(c, SPF_T_DEPOSIT);
    HIPCALL(c, deposit_currents(*R, c->stream, !filt));
  } else if (!filt && (rc = spf_currents_reset(c, run))) {
    return rc;
  }
  for (int32_t k = 0; k < n_lines; k++)
    if ((rc = spf_currents_add_line(c, run, line_dir[k], line_pos[k], line_start[k], line_end[k], line_J[k])))
      return rc;
  if (filt && (rc = spf_currents_scale(c, run, R->jm2))) return rc;   // src/integrate.c:177-178
  if (R->pc_set && (rc = spf_poisson_correct(c, run, (R->d.modes & SPF_SPACE_CHARGE) != 0, nullptr))) return rc;
  if ((rc = spf_fields_advance(c, run))) return rc;
  if ((rc = spf_particles_push(c, run, R->z_forces_start))) return rc;   // src/integrate.c:186-188
  for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)   // src/integrate.c:190-191
    if ((rc = cathode_current_density(c, run, k))) return rc;
  R->time_step++;
  return SPF_OK;
}


int spf_step_batch(spf_ctx *c, int32_t n_runs, const int32_t *run_ids) {
  if (!c || n_runs < 0 || (n_runs > 0 && !run_ids)) return fail(c, SPF_EINVAL, "bad batch arguments");
  std::vector<Run *> runs(n_runs);
  for (int32_t i = 0; i < n_runs; i++) {
    runs[i] = get_run(c, run_ids[i]);
    if (!runs[i]) return fail(c, SPF_EINVAL, "bad run id in the batch");
    for (int32_t j = 0; j < i; j++)
      if (runs[j] == runs[i]) return fail(c, SPF_EINVAL, "a run appears twice in the batch");
  }
  if (n_runs == 0) return SPF_OK;
  int rc;
  hipStream_t s = c->stream;
  PhaseTimer pt(c, SPF_T_STEP);
  // per run, in the order of src/integrate.c:134-183: the ramped imposed field,
  // the time, emission, J reset and deposition; the cell-sort runs with
  // particles (and without smearing, which takes the ordered deposition)
  // deposit together in one set of launches
  std::vector<Run *> together;
  for (int32_t i = 0; i < n_runs; i++) {
    Run *R = runs[i];
    if (R->ramp_time > 0 && (rc = spf_imposed_efield(c, run_ids[i], 0, R->ramp_time, R->flat_top_time))) return rc;
    R->time = R->time_step * R->d.dt + R->d.start_time;
    for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)
      if ((rc = cathode_emit_step(c, run_ids[i], k, nullptr))) return rc;
    if (R->cell_sort > 0 && R->time_step % R->cell_sort == 0 && (rc = spf_particles_sort(c, run_ids[i]))) return rc;
    if (R->jfm != 0) {
      // the J time filter (src/integrate.c:168-178): scaled, deposited on, scaled
      if ((rc = spf_currents_scale(c, run_ids[i], R->jm1))) return rc;
      if (R->d.modes & SPF_SPACE_CHARGE) {
        PhaseTimer dt(c, SPF_T_DEPOSIT);
        HIPCALL(c, deposit_currents(*R, s, false));
      }
      if ((rc = spf_currents_scale(c, run_ids[i], R->jm2))) return rc;
      continue;
    }
    if (R->d.modes & SPF_SPACE_CHARGE) {
      if (R->cell_sort > 0 && R->np_ub > 0 && (R->d.modes & (SPF_TM_FIELDS | SPF_TE_FIELDS)) &&
          deposit_slots_per_particle(R->d.modes) == 1) {
        together.push_back(R);
        continue;
      }
      PhaseTimer dt(c, SPF_T_DEPOSIT);
      HIPCALL(c, deposit_currents(*R, s, true));
    } else if ((rc = spf_currents_reset(c, run_ids[i]))) {
      return rc;
    }
  }
  if (together.size() == 1) {
    PhaseTimer dt(c, SPF_T_DEPOSIT);
    HIPCALL(c, deposit_currents(*together[0], s, true));
  } else if (!together.empty()) {
    PhaseTimer dt(c, SPF_T_DEPOSIT);
    hipError_t e = deposit_currents_multi(c, together.data(), (int)together.size());
    if (e == hipErrorInvalidValue) {   // the key spaces do not fit one 32-bit space: one run at a time
      for (Run *R : together) HIPCALL(c, deposit_currents(*R, s, true));
    } else {
      HIPCALL(c, e);
    }
  }
  // Poisson corrections: the runs due now solve together, grouped by accuracy and limit
  {
    std::vector<int32_t> solving;
    for (int32_t i = 0; i < n_runs; i++) {
      bool solve = false;
      if (runs[i]->pc_set &&
          (rc = poisson_pre(c, runs[i], (runs[i]->d.modes & SPF_SPACE_CHARGE) != 0, &solve)))
        return rc;
      if (solve) solving.push_back(i);
    }
    std::vector<char> done(solving.size(), 0);
    for (size_t g = 0; g < solving.size(); g++) {
      if (done[g]) continue;
      const spf_poisson_params &p = runs[solving[g]]->pc;
      std::vector<Run *> grp;
      std::vector<const double *> rhs;
      std::vector<size_t> idx;
      for (size_t h = g; h < solving.size(); h++) {
        const spf_poisson_params &q = runs[solving[h]]->pc;
        if (!done[h] && q.accuracy == p.accuracy && q.maximum_iterations == p.maximum_iterations) {
          done[h] = 1;
          grp.push_back(runs[solving[h]]);
          rhs.push_back(runs[solving[h]]->Q);
          idx.push_back(h);
        }
      }
      std::vector<int64_t> its(grp.size(), -1);
      {
        PhaseTimer ptp(c, SPF_T_POISSON);
        if ((rc = sor_solve_multi(c, grp.data(), rhs.data(), p.accuracy, p.maximum_iterations, (int)grp.size(),
                                  its.data())))
          return rc;
      }
      for (size_t k = 0; k < grp.size(); k++)
        if ((rc = poisson_post(c, grp[k], (grp[k]->d.modes & SPF_SPACE_CHARGE) != 0, its[k]))) return rc;
    }
  }
  // FDTD of every run in one ragged launch (plus its edge launch)
  {
    std::vector<FdtdArgs> fa;
    std::vector<Run *> fr;
    for (Run *R : runs)
      if (R->d.modes & (SPF_TM_FIELDS | SPF_TE_FIELDS)) {
        fa.emplace_back();
        fdtd_args(R, fa.back());
        fr.push_back(R);
      }
    if (!fa.empty()) {
      void *host, *dev;
      const size_t bytes = fdtd_multi_table_bytes((int)fa.size());
      if ((rc = batch_table(c, bytes, &host, &dev))) return rc;
      int64_t tb, eb;
      fdtd_multi_table(fa.data(), (int)fa.size(), host, &tb, &eb);
      PhaseTimer ft(c, SPF_T_FDTD);
      HIPCALL(c, batch_table_upload(c, host, dev, bytes));
      HIPCALL(c, launch_fdtd_multi(dev, (int)fa.size(), tb, eb, s));
      for (Run *R : fr) fdtd_flip(R);
    }
  }
  // push, event rows and compaction of every run in one ragged launch each
  {
    std::vector<PushArgs> pa(n_runs);
    std::vector<RowsArgs> ra(n_runs);
    std::vector<int64_t> nub(n_runs);
    for (int32_t i = 0; i < n_runs; i++) {
      push_args(runs[i], runs[i]->z_forces_start, pa[i], ra[i]);
      nub[i] = runs[i]->np_ub;
    }
    void *host, *dev;
    const size_t bytes = push_multi_table_bytes(n_runs);
    if ((rc = batch_table(c, bytes, &host, &dev))) return rc;
    const int64_t blocks = push_multi_table(pa.data(), nub.data(), ra.data(), n_runs, host);
    PhaseTimer ptp(c, SPF_T_PUSH);
    HIPCALL(c, batch_table_upload(c, host, dev, bytes));
    HIPCALL(c, launch_push_multi(dev, n_runs, blocks, s));
  }
  for (int32_t i = 0; i < n_runs; i++) {
    Run *R = runs[i];
    for (int32_t k = 0; k < (int32_t)R->cathodes.size(); k++)   // src/integrate.c:190-191
      if ((rc = cathode_current_density(c, run_ids[i], k))) return rc;
    R->time_step++;
  }
  return SPF_OK;
}

}  // extern "C"
