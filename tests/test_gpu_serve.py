"""The serving loop: requests admitted while others run, stepped together by
continuous batching, give each request the result it gets served alone."""
import pytest

from spiffe_amd import serve as V
from spiffe_amd.device import Context

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    reqs = V.request_stream(10, 40.0, seed=3, steps=12, dc_fraction=0.3)
    for r in reqs:   # small runs keep the test short
        if r["kind"] == "gun":
            r["cfg"] = dict(r["cfg"], dz=1.0e-4, particles=min(r["cfg"]["particles"], 300_000))
    assert any(r["kind"] == "dc" for r in reqs) and any(r["kind"] == "gun" for r in reqs)
    with Context(0) as ctx:
        mixed, st = V.serve_continuous(ctx, reqs, memory_budget=100e9)
        alone, _ = V.serve_one_at_a_time(ctx, reqs)
    assert st["peak_resident_runs"] > 1          # requests overlapped on the device
    for r in reqs:
        assert mixed[r["id"]]["digest"] == alone[r["id"]]["digest"], "request %d" % r["id"]
