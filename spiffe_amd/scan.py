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
arithmetic), which verify_batched() checks run by run.
"""
import math
import time

import numpy as np

from . import workloads as W
from .device import DeviceRun

SCAN_SEED = 20260
SCAN_POISSON = dict(step_interval=10, accuracy=1e-8, maximum_iterations=3000)


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


def build_run(ctx, cfg, grids=None):
    """A resident run of one scan configuration (grids caches geometry per mesh)."""
    key = cfg["dz"]
    if grids is not None and key in grids:
        g, f = grids[key]
    else:
        g = W.gun_grid(dz=cfg["dz"])
        f = W.gun_flags(g)
        if grids is not None:
            grids[key] = (g, f)
    run = DeviceRun(ctx, g, f, max_particles=cfg["particles"], screens=W.GUN_SCREENS, compat=False)
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


def one_at_a_time(ctx, cfgs, steps, grids=None):
    """Each run created, stepped `steps` times and destroyed before the next.
    Returns (device-resident stepping seconds, end-to-end seconds)."""
    step_s, t0 = 0.0, time.perf_counter()
    for cfg in cfgs:
        run = build_run(ctx, cfg, grids)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            run.step()
        ctx.synchronize()
        step_s += time.perf_counter() - t1
        run.close()
    return step_s, time.perf_counter() - t0


def batched(ctx, cfgs, steps, sub_batch, grids=None, keep=False):
    """All runs resident, each sub-batch of `sub_batch` runs stepped together
    `steps` times by spf_step_batch.  Returns (stepping seconds, end-to-end
    seconds, runs if keep else None)."""
    t0 = time.perf_counter()
    runs = [build_run(ctx, cfg, grids) for cfg in cfgs]
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
