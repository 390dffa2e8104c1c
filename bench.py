"""Benchmark of the per-step EM-PIC hot path on one GPU.

Workload: the production-size single run (1.6-cell S-band gun, dr = dz = 0.1 mm,
1051 x 421 grid lines, 1M macroparticles / 1 nC, TM+TE fields, Poisson
correction every 10 steps to 1e-8, one screen at z = 10 cm), FP64, synthetic
field and bunch contents (spiffe_amd.workloads).  A step is one pass of the
loop body of src/integrate.c:134-194: J reset, NGP deposition, Poisson
correction (every 10th step), FDTD, push with loss/screen handling.

Prints ONE JSON line (rank 0).  `value` is particle-pushes/s over the timed
steps with all state resident in HBM; cell-updates/s is reported beside it.
The `roofline` object is for the kernel that takes the most time per step,
timed with HIP events on the library's own stream; `cpu_baseline` is the
CPU restatement (oracle/) timed on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-pushes/s + cell-updates/s (% HBM roofline); scan runs/hr + p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per unit (FP64 = 8 B, flags 2 B)
B_CELL_TMTE = 146.0      # TM+TE cell-update: reads Ez Er Bphi Jz Jr Ephi Br Bz Jphi + flag, writes 9 arrays
B_PUSH = 196.0           # particle-push: 12 doubles + status read, 12 doubles written
B_SOR = 22.0             # SOR point of one half-sweep: Psi rhs flag read, half of points written
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE on this workload, scaled by
# the counters' reading for a plain 512 MiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); the
# push phase is its push_kernel (the rows kernel moves a few kB); None where not measured
TRAFFIC = {"sor_batch_kernel": 14.42e6, "fdtd_kernel": 66.04e6, "push phase": 187.68e6}
TIMER_EVERY = 8


class HipTimer:
    """HIP events recorded on an arbitrary stream (the library's), via the HIP runtime."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def time_ms(self, stream, fn, reps):
        a, b = self.event(), self.event()
        self.hip.hipEventRecord(a, stream)
        for _ in range(reps):
            fn()
        self.hip.hipEventRecord(b, stream)
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b)
        self.hip.hipEventDestroy(a)
        self.hip.hipEventDestroy(b)
        return ms.value


def build_run(ctx, n_particles):
    from spiffe_amd import workloads as W
    from spiffe_amd.device import DeviceRun
    g = W.gun_grid()
    f = W.gun_flags(g)
    run = DeviceRun(ctx, g, f, max_particles=n_particles, screens=W.GUN_SCREENS, compat=False)
    for k, v in W.gun_fields(g).items():
        run.upload(k, v)
    run.define_solenoid(**W.GUN_SOLENOID)
    run.append(W.gun_bunch(g, n_particles))
    run.setup_poisson_correction(**W.GUN_POISSON)
    return g, f, run


def kernels_at_scale(ctx, n_particles, dz, reps):
    """FDTD and push of the high-resolution single run's size (a 4201 x 1681-line
    gun at dz = dr = 25 um, 16M particles): device ms per launch from the
    library's HIP-event timers (every launch timed) and algorithmic bytes."""
    from spiffe_amd import workloads as W
    from spiffe_amd.device import DeviceRun
    g = W.gun_grid(dz=dz)
    f = W.gun_flags(g)
    run = DeviceRun(ctx, g, f, max_particles=n_particles, screens=W.GUN_SCREENS, compat=False)
    for k, v in W.gun_fields(g).items():
        run.upload(k, v)
    run.define_solenoid(**W.GUN_SOLENOID)
    run.append(W.gun_bunch(g, n_particles))
    for _ in range(2):
        run.advance_field_solutions()
        run.advance_particles()
    ctx.synchronize()
    np0, _ = run.count()
    ctx.timing(1)
    for _ in range(reps):
        run.advance_field_solutions()
    for _ in range(reps):
        run.advance_particles()
    tm = ctx.timings()
    ctx.timing(0)
    np1, _ = run.count()
    cells = (g.nz - 1) * (g.nr - 1)
    t_f = tm["fdtd"][0] / tm["fdtd"][1]
    t_p = tm["push"][0] / tm["push"][1]
    npk = 0.5 * (np0 + np1)
    out = {"grid_lines": [g.nz, g.nr], "cells": cells, "particles": npk, "launches": reps,
           "fdtd_ms": t_f, "fdtd_GBs": cells * B_CELL_TMTE / (t_f * 1e-3) / 1e9,
           "push_ms": t_p, "push_GBs": npk * B_PUSH / (t_p * 1e-3) / 1e9}
    out["fdtd_frac"] = out["fdtd_GBs"] / HBM_PEAK_GBS
    out["push_frac"] = out["push_GBs"] / HBM_PEAK_GBS
    run.close()
    return out


def cpu_baseline(g, f, n_particles, budget_s):
    """The restatement on one host core, same workload, bounded sample of steps."""
    from oracle.oracle import OracleSim
    from spiffe_amd import workloads as W
    o = OracleSim(g, f, max_particles=n_particles)
    for k, v in W.gun_fields(g).items():
        o.set_field(k, v)
    o.define_solenoid(**W.GUN_SOLENOID)
    o.append(W.gun_bunch(g, n_particles))
    o.setup_poisson_correction(**W.GUN_POISSON)
    steps, pushes = 0, 0
    t0 = time.perf_counter()
    while True:
        pushes += int(o.b.np)
        o.step(screens=W.GUN_SCREENS, compat=False)
        steps += 1
        if time.perf_counter() - t0 > budget_s and steps % 10 == 0:
            break
    dt = time.perf_counter() - t0
    return {"value": pushes / dt, "unit": "particle-pushes/s", "cores": 1, "kind": "port",
            "sample": "%d whole steps (incl. %d Poisson corrections) of the same 1M-particle gun workload, "
                      "%.1f s, serial restatement built -O2 -ffp-contract=off" % (steps, (steps + 9) // 10, dt)}



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-budget", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--at-scale", action="store_true",
                    help="also time FDTD and push at the high-resolution run's size (4201x1681 lines, 16M particles)")
    ap.add_argument("--scale-particles", type=int, default=16_000_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device copy (5 times), the reference for HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "copies": 5}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: independent gun runs shard across GPUs with no exchange, "
              "and this build measures the single-GPU path only (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from spiffe_amd import workloads as W
    from spiffe_amd.device import Context
    ctx = Context(0)
    g, f, run = build_run(ctx, args.particles)
    stream = ctx.stream()
    timer = HipTimer()
    cells = (g.nz - 1) * (g.nr - 1)

    for _ in range(args.warmup):
        run.step()
    ctx.synchronize()
    np0, _ = run.count()
    # timed region: K whole steps; the library's HIP-event timers on its stream sample one
    # call in TIMER_EVERY of each phase (event markers cost a few microseconds each)
    ctx.timing(TIMER_EVERY)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    tm = ctx.timings()
    ctx.timing(0)
    np1, _ = run.count()
    pushes = 0.5 * (np0 + np1) * args.steps
    ms_per_step = 1e3 * wall / args.steps

    def avg(name):   # mean device ms per launch over the sampled launches
        ms, n, _ = tm[name]
        return ms / n if n else None

    def per_step(name):   # sampled mean per launch x launches per step
        ms, n, launched = tm[name]
        return ms / n * launched / args.steps if n else 0.0

    t_fdtd, t_push, t_dep, t_sor = avg("fdtd"), avg("push"), avg("deposit"), avg("sor_batch")
    phase = {"fdtd_ms": t_fdtd, "push_ms": t_push, "deposit_ms": t_dep,
             "sor_batch_ms": t_sor, "sor_batches": tm["sor_batch"][2], "sor_batches_timed": tm["sor_batch"][1],
             "poisson_ms_per_correction": avg("poisson"), "poisson_corrections": tm["poisson"][2],
             "poisson_ms_per_step": per_step("poisson")}
    # device time per step spent in each kernel or phase
    share = {"fdtd_kernel": per_step("fdtd"), "push phase": per_step("push"), "deposit phase": per_step("deposit"),
             "sor_batch_kernel": per_step("sor_batch")}
    cells_all = (g.nz + 1) * (g.nr + 1)
    hs = ctx.lib.spf_sor_batch_half_sweeps()   # half-sweeps per SOR batch launch
    npk = 0.5 * (np0 + np1)
    algo = {  # algorithmic bytes per launch and measured ms per launch
        "fdtd_kernel": (cells * B_CELL_TMTE, t_fdtd, "%d cells x %d B (TM+TE cell-update)" % (cells, B_CELL_TMTE)),
        "sor_batch_kernel": (hs * cells_all * B_SOR, t_sor,
                             "%d half-sweeps x %d points x %d B (SOR point of one half-sweep)"
                             % (hs, cells_all, B_SOR)),
        "push phase": (npk * B_PUSH, t_push, "%d particles x %d B (particle-push)" % (npk, B_PUSH)),
    }
    dominant = max((k for k in share if k in algo and algo[k][1]), key=share.get)
    byts, ms, what = algo[dominant]
    ach = byts / (ms * 1e-3) / 1e9
    rl = {"bound": "hbm", "kernel": dominant, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": ach / HBM_PEAK_GBS, "traffic": TRAFFIC.get(dominant), "algorithmic_bytes_per_launch": byts,
          "ms_per_launch": ms, "units": what}
    gbs = {k: v[0] / (v[1] * 1e-3) / 1e9 for k, v in algo.items() if v[1]}
    # iteration count of one correction (outside the timed region)
    run.set_step(0)
    it = run.add_poisson_correction(True)

    scale = None
    if args.at_scale:
        run.close()
        run = None
        scale = kernels_at_scale(ctx, args.scale_particles, 25e-6, 20)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, f, args.particles, args.cpu_budget)

    out = {
        "metric": METRIC, "value": pushes / wall, "unit": "particle-pushes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "1.6-cell S-band gun, dr=dz=0.1 mm (1051x421 lines), 1M macroparticles 1 nC, "
                               "TM+TE, Gaussian solenoid 0.25 T at z=20 cm, Poisson every 10 steps to 1e-8, "
                               "screen z=10 cm",
                   "grid_lines": [g.nz, g.nr], "cells": cells, "particles": args.particles, "dt": g.dt},
        "cell_updates_per_s": cells * args.steps / wall,
        "phase_ms": phase, "poisson_iterations": it, "kernel_ms_per_step": share,
        "hbm_frac": {k: v / HBM_PEAK_GBS for k, v in gbs.items()},
        "roofline": rl, "cpu_baseline": cpu,
    }
    if scale:
        out["kernels_at_scale"] = scale
    print(json.dumps(out))
    if run:
        run.close()
    ctx.close()


if __name__ == "__main__":
    main()
