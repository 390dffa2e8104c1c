"""Design scans of the 1.6-cell gun: many independent runs on one GPU.

scan_configs() draws the 64-run scan (seeded): mesh dz = dr ~ U(0.05, 0.1) mm,
macroparticles ~ logU(1e5, 4e6), RF phase ~ U(0, 60 deg), peak gradient ~
U(80, 140) MV/m, solenoid peak ~ U(0.15, 0.35) T, charge ~ logU(10 pC, 2 nC),
cathode radius ~ U(0.5, 2) mm.  Each run is the gun workload of
spiffe_amd.workloads with those values: the synthetic standing-wave field
scaled to the gradient and by cos(phase) (the field the bunch sees at its
phase), the Gaussian solenoid at the drawn peak, a Gaussian bunch of the drawn
count, charge and radius, Poisson correction every 10 steps to 1e-8 capped at
3000 half-sweeps, a screen at z = 10 cm.

Two ways to run a scan on one GPU:
  one_at_a_time: create a run, step it, destroy it, the next;
  batched: all runs resident, stepped together by spf_step_batch (ragged
  launches over the runs), in sub-batches whose size tune_batch() picks by
  timing the candidates on the scan's own runs.
Both give every run the same bits (spf_step_batch keeps each run's
arithmetic), which the bench checks run by run.

Scan runs are cell-sort runs (particles re-sorted by cell, deposition in the
library's own fixed order, results within the design tolerances of the
reference's); tune_cadence() picks the re-sort interval from one timed probe.
"""
import math
import time

import numpy as np

from . import workloads as W
from .device import DeviceRun

SCAN_SEED = 20260
SCAN_POISSON = dict(step_interval=10, accuracy=1e-8, maximum_iterations=3000)
SCAN_CELL_SORT = 10   # re-sort interval until tune_cadence has chosen one


def scan_configs(n=64, seed=SCAN_SEED):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        out.append({
            "index": k,
            "dz": float(rng.uniform(0.05e-3, 0.1e-3)),
            "particles": int(round(math.exp(rng.uniform(math.log(1e5), math.log(4e6))))),
            "phase_deg": float(rng.uniform(0, 60)),
            "gradient": float(rng.uniform(80e6, 140e6)),
            "solenoid": float(rng.uniform(0.15, 0.35)),
            "charge": float(math.exp(rng.uniform(math.log(10e-12), math.log(2e-9)))),
            "radius": float(rng.uniform(0.5e-3, 2e-3)),
            "seed": int(rng.integers(1, 2 ** 31 - 1)),
        })
    return out


def build_run(ctx, cfg, grids=None, cell_sort=SCAN_CELL_SORT):
    """A resident run of one scan configuration (grids caches geometry per mesh)."""
    key = cfg["dz"]
    if grids is not None and key in grids:
        g, f = grids[key]
    else:
        g = W.gun_grid(dz=cfg["dz"])
        f = W.gun_flags(g)
        if grids is not None:
            grids[key] = (g, f)
    run = DeviceRun(ctx, g, f, max_particles=cfg["particles"], screens=W.GUN_SCREENS, compat=False,
                    cell_sort=cell_sort)
    ez = W.gun_fields(g, E0=cfg["gradient"])["Ez"] * math.cos(math.radians(cfg["phase_deg"]))
    run.upload("Ez", ez)
    run.define_solenoid(**dict(W.GUN_SOLENOID, Bz_peak=cfg["solenoid"]))
    run.append(W.gun_bunch_gaussian(g, cfg["particles"], charge=cfg["charge"], radius=cfg["radius"],
                                    seed=cfg["seed"]))
    run.setup_poisson_correction(**SCAN_POISSON)
    return run


def footprint_bytes(ctx, cfg):
    run = build_run(ctx, cfg)
    b = run.footprint()
    run.close()
    return b


def one_at_a_time(ctx, cfgs, steps, grids=None, cell_sort=SCAN_CELL_SORT):
    """Each run created, stepped `steps` times and destroyed before the next.
    Returns (device-resident stepping seconds, end-to-end seconds)."""
    step_s, t0 = 0.0, time.perf_counter()
    for cfg in cfgs:
        run = build_run(ctx, cfg, grids, cell_sort)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            run.step()
        ctx.synchronize()
        step_s += time.perf_counter() - t1
        run.close()
    return step_s, time.perf_counter() - t0


def batched(ctx, cfgs, steps, sub_batch, grids=None, keep=False, cell_sort=SCAN_CELL_SORT):
    """All runs resident, each sub-batch of `sub_batch` runs stepped together
    `steps` times by spf_step_batch.  Returns (stepping seconds, end-to-end
    seconds, runs if keep else None)."""
    t0 = time.perf_counter()
    runs = [build_run(ctx, cfg, grids, cell_sort) for cfg in cfgs]
    ctx.synchronize()
    t1 = time.perf_counter()
    for i in range(0, len(runs), sub_batch):
        part = runs[i:i + sub_batch]
        for _ in range(steps):
            ctx.step_batch(part)
    ctx.synchronize()
    step_s = time.perf_counter() - t1
    if not keep:
        for r in runs:
            r.close()
        runs = None
    return step_s, time.perf_counter() - t0, runs


def tune_batch(ctx, runs, candidates=(8, 16, 32, 64), steps=3):
    """The autotuner: times `steps` batched steps of the given resident runs
    for each sub-batch size (the steps advance the runs: every size gives the
    same bits) and returns the fastest size and the timings.  The probe steps
    include the scan's first Poisson correction when it falls in them, so the
    sizes are compared on one step mix each: the first probe step of every
    candidate is the same kind of step."""
    timings = {}
    for s in candidates:
        if s > len(runs) and s != candidates[0]:
            continue
        ctx.synchronize()
        t0 = time.perf_counter()
        for i in range(0, len(runs), s):
            for _ in range(steps):
                ctx.step_batch(runs[i:i + s])
        ctx.synchronize()
        timings[s] = time.perf_counter() - t0
    best = min(timings, key=timings.get)
    return best, timings


def tune_cadence(ctx, runs, probe_steps=20, candidates=(2, 4, 5, 8, 10, 16, 20)):
    """The re-sort interval for a batch of resident cell-sort runs: one re-sort
    of every run is timed, then `probe_steps` batched steps with no re-sort in
    between, each timed (a step with a Poisson correction, which costs the same
    at any interval, counts as the median of the others); the interval N
    chosen minimises the mean step time (re-sort + steps 1..N after it) / N
    over the candidates (the per-step time grows as particles drift from the
    cell order).  Returns (N, probe timings).  The probe steps advance the
    runs."""
    for r in runs:
        r.set_cell_sort(1 << 40)   # no re-sort inside the probe
    ctx.synchronize()
    t0 = time.perf_counter()
    for r in runs:
        r.sort_particles()
    ctx.synchronize()
    t_sort = time.perf_counter() - t0
    t_steps, corrected = [], []
    for _ in range(probe_steps):
        solves = runs[0].poisson_stats()[1]
        t0 = time.perf_counter()
        ctx.step_batch(runs)
        ctx.synchronize()
        t_steps.append(time.perf_counter() - t0)
        corrected.append(runs[0].poisson_stats()[1] != solves)
    # a step with a Poisson correction costs the same whatever the interval:
    # it stands in with the median of the other steps
    plain = sorted(t for t, c in zip(t_steps, corrected) if not c)
    if plain:
        t_steps = [plain[len(plain) // 2] if c else t for t, c in zip(t_steps, corrected)]
    cost = {}
    for n in candidates:
        if n <= probe_steps:
            cost[n] = (t_sort + sum(t_steps[:n])) / n
    best = min(cost, key=cost.get)
    for r in runs:
        r.set_cell_sort(best)
    return best, {"sort_s": t_sort, "step_s": t_steps, "mean_step_s": cost}


def tune_sor_shape(ctx, runs, shapes=(0, 1)):
    """The SOR tile shape for a batch of resident runs: for each shape, the
    batch is stepped up to and through its next Poisson correction and the
    correction step is timed (every scan correction runs to the same
    half-sweep cap, so the steps compare).  The context keeps the fastest
    shape.  Psi and every other result are the same for either shape; the
    probe steps advance the runs.  Returns (shape, correction step seconds)."""
    times = {}
    for shape in shapes:
        ctx.set_sor_shape(shape)
        for _ in range(SCAN_POISSON["step_interval"] + 1):
            solves = runs[0].poisson_stats()[1]
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.step_batch(runs)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if runs[0].poisson_stats()[1] != solves:
                times[shape] = dt
                break
    best = min(times, key=times.get)
    ctx.set_sor_shape(best)
    return best, times
