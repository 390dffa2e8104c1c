"""Benchmark of the per-step EM-PIC hot path on one GPU.

Default workload: the high-resolution single run -- the 1.6-cell S-band gun at
dr = dz = 25 um (4201 x 1681 grid lines, 7.06M cells), 16M macroparticles
carrying 100 pC (Gaussian, sigma_t 3 ps, radius 0.6 mm), TM+TE fields, a
Gaussian solenoid (0.25 T at z = 20 cm), Poisson correction every 5 steps to a
residual of 1e-8 (cap 40 000 half-sweeps), one screen at z = 10 cm, FP64,
synthetic field and bunch contents (spiffe_amd.workloads).  `--workload gun`
runs the production-size run instead (1051 x 421 lines, 1M particles,
correction every 10 steps capped at 3000 half-sweeps).

A step is one pass of the loop body of src/integrate.c:134-194: J reset, NGP
deposition, Poisson correction (every 5th step), FDTD, push with
loss/screen handling.  Prints ONE JSON line.  `value` is particle-pushes/s over
the timed steps with all state resident in HBM; cell-updates/s is reported
beside it.  The `roofline` object is for the kernel with the most device time
per step, timed with HIP events on the library's own stream; `cpu_baseline`
is the CPU restatement (oracle/) on the host cores, from a bounded sample of
the same workload.
"""
import argparse
import ctypes
import json
import os
import platform
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
B_DEPOSIT = 52.0         # deposit, per particle: z r vz vr vphi Q + status
B_SOR = 22.0             # SOR point of one half-sweep: Psi rhs flag read, half of points written
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE on each workload, scaled by the
# counters' reading for a plain 512 MiB device copy (profiles/*_hbm_traffic.md); None where not measured
TRAFFIC = {
    "hires": {"sor_batch_kernel": 195.20e6, "fdtd_kernel": 1086.81e6, "push phase": 3016.52e6},
    "gun": {"sor_batch_kernel": 32.56e6, "fdtd_kernel": 66.03e6, "push phase": 188.26e6},
}
TIMER_EVERY = 2


def build_run(ctx, workload, n_particles, cell_sort=0, events=False):
    from spiffe_amd import workloads as W
    from spiffe_amd.device import DeviceRun
    screens = W.GUN_SCREENS
    if workload == "hires":
        g = W.gun_grid(dz=W.HIRES_DZ)
        bunch = W.gun_bunch_gaussian(g, n_particles, charge=W.HIRES_CHARGE, radius=W.HIRES_RADIUS)
        pc = W.HIRES_POISSON
        if events:   # a screen inside the bunch, and particles leaving through the cathode plane and the wall
            screens = W.HIRES_EVENT_SCREENS
            lv = W.gun_leavers(g, 100_000)
            bunch = {k: np.concatenate([bunch[k], lv[k]]) for k in bunch}
    else:
        g = W.gun_grid()
        bunch = W.gun_bunch(g, n_particles)
        pc = W.GUN_POISSON
    f = W.gun_flags(g)
    run = DeviceRun(ctx, g, f, max_particles=len(bunch["z"]), screens=screens, compat=False, cell_sort=cell_sort)
    for k, v in W.gun_fields(g).items():
        run.upload(k, v)
    run.define_solenoid(**W.GUN_SOLENOID)
    run.append(bunch)
    run.setup_poisson_correction(**pc)
    return g, f, bunch, pc, run


def host_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc) or nproc)
    return model, nproc, threads


def cpu_phase_sample(g, f, bunch, pc, threaded, sor_iterations):
    """The restatement on the same workload, phase by phase: one FDTD step,
    one J reset + deposition, one push of every particle, a fixed count of SOR
    half-sweeps and one capped correction (its Gauss-law charge, charge
    removal, field update and recomputed charge); the step time is assembled
    as FDTD + deposit + push + (correction extras + iterations x half-sweep) /
    interval, with the iteration count the GPU measured on this workload."""
    from oracle.oracle import OracleSim
    from spiffe_amd import workloads as W
    n = len(bunch["z"])
    o = OracleSim(g, f, max_particles=n, threaded=threaded, has_solenoid=True)
    for k, v in W.gun_fields(g).items():
        o.set_field(k, v)
    o.define_solenoid(**W.GUN_SOLENOID)
    o.append(bunch)
    t = {}
    t0 = time.perf_counter()
    o.zero_current_arrays()
    o.add_particle_currents()
    t["deposit"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.advance_field_solutions()
    t["fdtd"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.advance_particles(W.GUN_SCREENS, False)
    t["push"] = time.perf_counter() - t0
    hs = 8
    o.set_field("Q", np.full(g.shape, 1e3))   # a right-hand side the 8 half-sweeps cannot solve
    t0 = time.perf_counter()
    o.solve_poisson_cyl("Psi", "Q", 1e-300, hs)
    t["sor_half_sweep"] = (time.perf_counter() - t0) / hs
    o.setup_poisson_correction(**dict(pc, maximum_iterations=2))
    o.f.time_step = 0
    t0 = time.perf_counter()
    o.add_poisson_correction(True)
    t["correction_extras"] = max(time.perf_counter() - t0 - 2 * t["sor_half_sweep"], 0.0)
    step = t["fdtd"] + t["deposit"] + t["push"] + \
        (t["correction_extras"] + sor_iterations * t["sor_half_sweep"]) / pc["step_interval"]
    del o
    return n / step, step, t


def scan_bench(n_runs, steps):
    """The design scan: n_runs gun runs of the seeded scan distribution, each
    stepped `steps` times (corrections at steps 0, 10, ...), one at a time and
    batched; runs/hour of each from the stepping time with state resident, and
    end to end (set-up included).  Four runs are checked bit for bit between
    the two."""
    from spiffe_amd import scan as S
    from spiffe_amd.device import Context
    cfgs = S.scan_configs(n_runs)
    out = {"metric": "scan runs/hr (batched vs one-at-a-time, one GPU)", "unit": "runs/hr", "n_gpus": 1,
           "runs": n_runs, "steps_per_run": steps, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "64-run 1.6-cell gun scan: dz=dr U(0.05,0.1) mm, particles logU(1e5,4e6), "
                                  "phase U(0,60) deg, gradient U(80,140) MV/m, solenoid U(0.15,0.35) T, "
                                  "charge logU(10 pC,2 nC), cathode radius U(0.5,2) mm; Poisson every 10 steps "
                                  "to 1e-8 capped at 3000 half-sweeps; particles re-sorted by cell at the "
                                  "interval the autotuner picks (the same for both ways of running)",
                      "seed": S.SCAN_SEED}}
    with Context(0) as ctx:
        grids = {}
        out["total_particles"] = int(sum(c["particles"] for c in cfgs))
        # the autotuner, on the scan's own resident runs (steps between corrections):
        # the sub-batch size, then the re-sort interval
        runs = [S.build_run(ctx, c, grids) for c in cfgs]
        ctx.step_batch(runs)
        best, timings = S.tune_batch(ctx, runs, steps=2)
        cadence, cad_probe = S.tune_cadence(ctx, runs)
        shape, shape_probe = S.tune_sor_shape(ctx, runs)
        for r in runs:
            r.close()
        # one at a time: each run solves with the shape its own grid size picks
        ctx.set_sor_shape(-1)
        seq_step, seq_all = S.one_at_a_time(ctx, cfgs, steps, grids, cell_sort=cadence)
        ctx.set_sor_shape(shape)
        bat_step, bat_all, kept = S.batched(ctx, cfgs, steps, best, grids, keep=True, cell_sort=cadence)
        # bit-for-bit check of four runs: batched state against a run stepped alone
        checked = []
        for i in (0, n_runs // 3, 2 * n_runs // 3, n_runs - 1):
            alone = S.build_run(ctx, cfgs[i], grids, cell_sort=cadence)
            for _ in range(steps):
                alone.step()
            same = all(np.array_equal(alone.download(k).view(np.uint64), kept[i].download(k).view(np.uint64))
                       for k in ("Ez", "Er", "Bphi", "Psi", "Jz"))
            pa, pb = alone.particles(), kept[i].particles()
            same = same and all(np.array_equal(pa[k].view(np.uint64), pb[k].view(np.uint64))
                                for k in ("z", "r", "pz", "pr"))
            same = same and alone.poisson_stats() == kept[i].poisson_stats()
            checked.append({"run": i, "bitwise_equal": bool(same)})
            alone.close()
        for r in kept:
            r.close()
    out.update({"value": n_runs / bat_step * 3600, "one_at_a_time_runs_per_hr": n_runs / seq_step * 3600,
                "speedup": seq_step / bat_step, "batched_s": bat_step, "one_at_a_time_s": seq_step,
                "end_to_end": {"batched_runs_per_hr": n_runs / bat_all * 3600,
                               "one_at_a_time_runs_per_hr": n_runs / seq_all * 3600,
                               "batched_s": bat_all, "one_at_a_time_s": seq_all},
                "autotune": {"sub_batch": best, "probe_s": timings, "cell_sort_interval": cadence,
                             "sor_shape": shape, "sor_shape_probe_s": shape_probe,
                             "cell_sort_probe": {"sort_s": cad_probe["sort_s"],
                                                 "mean_step_s": cad_probe["mean_step_s"]}},
                "bitwise_checks": checked})
    return out


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return 288e9
    return float(free.value)


def serve_bench(n_requests, steps):
    """The request stream, time-compressed: each request is `steps` whole steps
    of its run (a full run to the screen is about 50x longer), so arrivals come
    50x faster than the configured 20 runs/min to keep the load the GPU sees.
    Both schedulers serve the same stream; every request's result digest is
    compared between them."""
    from spiffe_amd import serve as V
    from spiffe_amd.device import Context
    compression = 2000.0 / steps
    rate = 20.0 / 60.0 * compression
    reqs = V.request_stream(n_requests, rate, steps=steps)
    with Context(0) as ctx:
        budget = 0.9 * device_free_bytes()
        done_c, st_c = V.serve_continuous(ctx, reqs, budget)
        print("continuous batching served %d requests" % len(reqs), file=sys.stderr, flush=True)
        done_1, st_1 = V.serve_one_at_a_time(ctx, reqs)
        print("one-at-a-time queue served %d requests" % len(reqs), file=sys.stderr, flush=True)
        differ = [r["id"] for r in reqs if done_c[r["id"]]["digest"] != done_1[r["id"]]["digest"]]
        serial = V.serve_serial(ctx, reqs, differ)
    same = len(reqs) - len(differ)
    mismatches = [{"id": i, "kind": reqs[i]["kind"], "cfg": reqs[i]["cfg"],
                   "continuous_is_serial": done_c[i]["digest"] == serial[i],
                   "one_at_a_time_is_serial": done_1[i]["digest"] == serial[i]} for i in differ]
    return {"metric": "serving runs/hr + p50/p99 time-to-result (continuous batching vs one-at-a-time queue)",
            "value": st_c["runs_per_hr"], "unit": "runs/hr", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "%d requests, Poisson arrivals at 20 runs/min time-compressed %gx (%.2f/s); "
                                   "80%% 1.6-cell gun runs from the scan distribution, 20%% DC-gun runs (100 kV, "
                                   "2 cm gap, 3 mm cathode, 2 mm bore, 0.05 mm mesh, 200k macroparticles 50 pC "
                                   "10 ps flat-top emitted); %d steps per request"
                                   % (n_requests, compression, rate, steps),
                       "memory_budget_bytes": budget, "dc_requests": sum(r["kind"] == "dc" for r in reqs)},
            "continuous": st_c, "one_at_a_time": st_1,
            "p50_s": st_c["p50_s"], "p99_s": st_c["p99_s"],
            "results_identical": "%d/%d" % (same, len(reqs)), "mismatches": mismatches}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["hires", "gun", "scan", "serve"], default="hires",
                    help="hires: the high-resolution single run (the headline); gun: the production-size run; "
                         "scan: the 64-run design scan, batched against one at a time; serve: the request "
                         "stream, continuous batching against a one-at-a-time queue")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-steps", type=int, default=40)
    ap.add_argument("--scan-runs", type=int, default=64)
    ap.add_argument("--scan-steps", type=int, default=40)
    ap.add_argument("--particles", type=int, default=0, help="0: the workload's own count")
    ap.add_argument("--cell-sort", type=int, default=-1,
                    help="re-sort particles by cell every N steps (deposition then sums in a fixed order of its "
                         "own); 0 keeps the reference's particle order and sums; default: 20 for hires (its "
                         "configuration), 0 for gun")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--events", action="store_true",
                    help="hires only: add a screen inside the bunch (about 110k crossings a step) and 100k particles "
                         "that leave through the cathode plane and the wall, so the event rows and the compaction are "
                         "timed in every step")
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

    if args.workload == "scan":
        print(json.dumps(scan_bench(args.scan_runs, args.scan_steps)))
        return
    if args.workload == "serve":
        print(json.dumps(serve_bench(args.serve_requests, args.serve_steps)))
        return

    from spiffe_amd import workloads as W
    from spiffe_amd.device import Context
    n_particles = args.particles or (W.HIRES_PARTICLES if args.workload == "hires" else 1_000_000)
    cell_sort = args.cell_sort if args.cell_sort >= 0 else (20 if args.workload == "hires" else 0)
    ctx = Context(0)
    g, f, bunch, pc, run = build_run(ctx, args.workload, n_particles, cell_sort, args.events and args.workload == "hires")
    cells = (g.nz - 1) * (g.nr - 1)

    for _ in range(args.warmup):
        run.step()
    ctx.synchronize()
    ring = 300_000 * (args.steps + args.warmup) if args.events else 1 << 20
    if args.events:
        run.drain(max_rows=ring, max_lost=ring)   # the warm-up steps' rows
    np0, _ = run.count()
    _, solves0, hs0 = run.poisson_stats()
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
    if args.events:   # the timed steps' rows, drained after the clock stopped
        rows, lost = run.drain(max_rows=ring, max_lost=ring)
        srows0, lost0, srows1, lost1 = 0, 0, len(rows), len(lost)
        del rows, lost
    last_it, solves1, hs1 = run.poisson_stats()
    pushes = 0.5 * (np0 + np1) * args.steps
    ms_per_step = 1e3 * wall / args.steps

    def avg(name):   # mean device ms per launch over the sampled launches
        ms, n, _ = tm[name]
        return ms / n if n else None

    def per_step(name):   # sampled mean per launch x launches per step
        ms, n, launched = tm[name]
        return ms / n * launched / args.steps if n else 0.0

    hs = run.sor_batch_half_sweeps()   # half-sweeps per SOR batch launch of this run's tile shape
    corrections = solves1 - solves0
    sor_half_sweeps = hs1 - hs0
    t_fdtd, t_push, t_dep, t_sor = avg("fdtd"), avg("push"), avg("deposit"), avg("sor_batch")
    phase = {"fdtd_ms": t_fdtd, "push_ms": t_push, "deposit_ms": t_dep,
             "sor_batch_ms": t_sor, "sor_batches": tm["sor_batch"][2], "sor_batches_timed": tm["sor_batch"][1],
             "cell_sort_ms": avg("cell_sort"), "cell_sorts": tm["cell_sort"][2],
             "poisson_ms_per_correction": avg("poisson"), "poisson_corrections": corrections,
             "poisson_ms_per_step": per_step("poisson")}
    # device time per step spent in each kernel or phase
    share = {"fdtd_kernel": per_step("fdtd"), "push phase": per_step("push"), "deposit phase": per_step("deposit"),
             "sor_batch_kernel": per_step("sor_batch"), "cell sort": per_step("cell_sort")}
    points = (g.nz + 1) * (g.nr + 1)
    npk = 0.5 * (np0 + np1)
    algo = {  # algorithmic bytes per launch and measured ms per launch
        "fdtd_kernel": (cells * B_CELL_TMTE, t_fdtd, "%d cells x %d B (TM+TE cell-update)" % (cells, B_CELL_TMTE)),
        "sor_batch_kernel": (hs * points * B_SOR, t_sor,
                             "%d half-sweeps x %d points x %d B (SOR point of one half-sweep)"
                             % (hs, points, B_SOR)),
        "push phase": (npk * B_PUSH, t_push, "%d particles x %d B (particle-push)" % (npk, B_PUSH)),
        "deposit phase": (npk * B_DEPOSIT, t_dep, "%d particles x %d B (deposit)" % (npk, B_DEPOSIT)),
    }
    dominant = max((k for k in share if k in algo and algo[k][1]), key=share.get)
    byts, ms, what = algo[dominant]
    ach = byts / (ms * 1e-3) / 1e9
    rl = {"bound": "hbm", "kernel": dominant, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": ach / HBM_PEAK_GBS, "traffic": TRAFFIC[args.workload].get(dominant),
          "algorithmic_bytes_per_launch": byts, "ms_per_launch": ms, "units": what}
    gbs = {k: v[0] / (v[1] * 1e-3) / 1e9 for k, v in algo.items() if v[1]}
    iters_per_correction = sor_half_sweeps / corrections if corrections else 0.0
    converged = last_it >= 0
    run.close()
    ctx.close()

    cpu = None
    if not args.no_cpu:
        model, nproc, threads = host_info()
        its = iters_per_correction or pc["maximum_iterations"]
        v_ser, step_ser, t_ser = cpu_phase_sample(g, f, bunch, pc, False, its)
        v_omp, step_omp, t_omp = cpu_phase_sample(g, f, bunch, pc, True, its)
        what_sample = ("per-phase sample of the same %d-particle workload: 1 FDTD step, 1 deposit, 1 push of every "
                       "particle, 8 SOR half-sweeps and 1 correction capped at 2 half-sweeps; step = FDTD + deposit "
                       "+ push + (correction extras + %.0f half-sweeps) / %d, the GPU's mean iterations per "
                       "correction; restatement built -O2 -ffp-contract=off" % (npk, its, pc["step_interval"]))
        cpu = {"value": v_omp, "unit": "particle-pushes/s", "cores": threads, "kind": "port",
               "sample": what_sample + ", OpenMP over %d threads" % threads,
               "cpu_model": model, "nproc": nproc, "s_per_step": step_omp,
               "phase_s": {k: round(v, 4) for k, v in t_omp.items()},
               "serial": {"value": v_ser, "cores": 1, "s_per_step": step_ser,
                          "phase_s": {k: round(v, 4) for k, v in t_ser.items()}}}

    if args.workload == "hires":
        wl = ("1.6-cell S-band gun, dr=dz=25 um (4201x1681 lines, 7.06M cells), 16M macroparticles 100 pC "
              "(Gaussian sigma_t 3 ps, r 0.6 mm), TM+TE, Gaussian solenoid 0.25 T at z=20 cm, Poisson correction "
              "every 5 steps to residual 1e-8 (cap 40000 half-sweeps), screen z=10 cm")
    else:
        wl = ("1.6-cell S-band gun, dr=dz=0.1 mm (1051x421 lines), 1M macroparticles 1 nC, TM+TE, Gaussian "
              "solenoid 0.25 T at z=20 cm, Poisson correction every 10 steps capped at 3000 half-sweeps "
              "(accuracy 1e-8 not reached), screen z=10 cm")
    out = {
        "metric": METRIC, "value": pushes / wall, "unit": "particle-pushes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "grid_lines": [g.nz, g.nr], "cells": cells, "particles": n_particles,
                   "dt": g.dt, "poisson": dict(pc),
                   "cell_sort": ("particles re-sorted by half-cell every %d steps; deposition sums per cell in a fixed "
                                 "order of its own (wavefront tree sums folded in storage order)" % cell_sort
                                 if cell_sort else
                                 "none: particles in the reference's order, deposition in the reference's order")},
        "cell_updates_per_s": cells * args.steps / wall,
        "events": ({"screen_rows_per_step": (srows1 - srows0) / args.steps, "lost_per_step": (lost1 - lost0) / args.steps}
                   if args.events else None),
        "phase_ms": phase, "sor_half_sweeps_per_correction": iters_per_correction,
        "poisson_last_converged": converged,
        "kernel_ms_per_step": share,
        "hbm_frac": {k: v / HBM_PEAK_GBS for k, v in gbs.items()},
        "roofline": rl, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
