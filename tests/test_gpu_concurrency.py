"""A run's bits do not depend on what else runs on the GPU: batched steps of
three gun runs (one of them solving to its half-sweep limit) give the same
fields, particles and screen rows whether or not another context, on another
thread, is building, stepping and reading back runs of its own at the same
time.  (Kernels of other queues change which workgroups share a CU and how
their wavefronts interleave; a workgroup that read LDS before the barrier
that publishes it gave different bits under such load.)"""
import threading
import time

import numpy as np
import pytest

from spiffe_amd import serve as V
from spiffe_amd.abi import FIELDS
from spiffe_amd.device import Context

pytestmark = pytest.mark.gpu

REQS = V.request_stream(64, 10.0, seed=7, steps=12)
GUNS = [r for r in REQS if r["kind"] == "gun" and 150_000 <= r["cfg"]["particles"] <= 700_000][:3]


def run_state(d):
    out = {n: d.download(n).view(np.uint64).copy() for n in FIELDS}
    for k, v in d.particles().items():
        out["p_" + k] = np.ascontiguousarray(v).view(np.uint8).copy()
    out["rows"] = d.drain()[0].view(np.uint8).copy()
    out["stats"] = np.array(d.poisson_stats())
    return out


def background(stop, done):
    with Context(0) as b:
        k = 0
        while not stop.is_set():
            d = V.build(b, V.prepare(REQS[(5 * k + 1) % len(REQS)], {}))
            for _ in range(3):
                d.step()
            V.digest(d, [d.drain()[0]])
            d.close()
            k += 1
            done[0] = k


def batched(ctx, steps, load):
    stop, done = threading.Event(), [0]
    th = threading.Thread(target=background, args=(stop, done)) if load else None
    if th:
        th.start()
        while done[0] == 0 and th.is_alive():   # the other context is busy before we start
            time.sleep(0.01)
    try:
        runs = [V.build(ctx, V.prepare(r, {})) for r in GUNS]
        for _ in range(steps):
            ctx.step_batch(runs)
        out = [run_state(d) for d in runs]
        for d in runs:
            d.close()
    finally:
        stop.set()
        if th:
            th.join()
    return out


def test_batched_bits_unchanged_by_another_contexts_work():
    assert len(GUNS) == 3
    with Context(0) as ctx:
        ref = batched(ctx, 12, False)
        assert any(s["stats"][0] == -1 for s in ref), "one solve should run to its limit"
        for trial in range(3):
            got = batched(ctx, 12, True)
            for i, (a, b) in enumerate(zip(ref, got)):
                for k in a:
                    assert np.array_equal(a[k], b[k]), "trial %d run %d %s" % (trial, i, k)
