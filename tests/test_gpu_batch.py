"""Many runs stepped together (spf_step_batch) give each run the bits it gets
stepped alone (spf_step): ragged SOR, FDTD, push and rows launches over runs
of different grids, particle counts, Poisson settings and modes."""
import numpy as np
import pytest

from _support import TMTE, bits, gun_flags, gun_grid, pillbox_flags, pillbox_grid, random_fields, random_particles
from spiffe_amd import abi
from spiffe_amd.device import Context, DeviceRun

pytestmark = pytest.mark.gpu

NAMES = ["Ez", "Er", "Bphi", "Bphi2", "Ephi", "Br", "Br2", "Bz", "Bz2", "Jz", "Jr", "Jphi", "Psi", "Q"]


def scenarios():
    """(grid, flags, particles, poisson settings, compat, cathode) of six different runs."""
    out = []
    for k, (nz, nr, n, pc, compat) in enumerate([
            (141, 61, 4000, dict(step_interval=3, accuracy=1e-5, maximum_iterations=400), False),
            (101, 45, 2500, dict(step_interval=3, accuracy=1e-5, maximum_iterations=400), True),
            (181, 81, 6000, dict(step_interval=4, accuracy=1e-6, maximum_iterations=600), False),
            (121, 51, 3000, dict(step_interval=3, accuracy=1e-5, maximum_iterations=400,
                                 error_charge_threshold=-1e-3), False),
            (161, 71, 5000, None, False)]):
        g = gun_grid(nz=nz, nr=nr, modes=TMTE | abi.SPACE_CHARGE)
        f = gun_flags(g)
        out.append((g, f, random_particles(g, n, 300 + k, flags=f, pz_scale=1.0, q=-1e-13), pc, compat, None,
                    random_fields(g, 400 + k, scale=2e6)))
    g = pillbox_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    cathode = dict(z_position=0.0, outer_radius=1e-3, current_density=1e4, electrons_per_macroparticle=1e3,
                   number_per_step=0, start_time=0.0, stop_time=20e-12, initial_pz=2e-3, random_number_seed=12345)
    out.append((g, pillbox_flags(g), None, dict(step_interval=3, accuracy=1e-5, maximum_iterations=300), False,
                cathode, random_fields(g, 410, names=["Ez"], scale=1e6)))
    return out


def build(ctx, sc):
    g, f, parts, pc, compat, cathode, fields = sc
    cap = len(parts["z"]) if parts is not None else 20000
    d = DeviceRun(ctx, g, f, max_particles=cap, screens=[(g.zmin + 0.4 * (g.zmax - g.zmin), 1)], compat=compat)
    for k, v in fields.items():
        d.upload(k, v)
    if parts is not None:
        d.append(parts)
    if pc:
        d.setup_poisson_correction(**pc)
    if cathode:
        d.define_cathode(**cathode)
    return d


def state(d):
    out = {n: d.download(n) for n in NAMES}
    out.update({"p_" + k: v for k, v in d.particles().items()})
    out["stats"] = d.poisson_stats()
    return out


def test_batched_steps_match_single_steps():
    steps = 13
    with Context(0) as ctx:
        alone, alone_rows = [], []
        for sc in scenarios():
            d = build(ctx, sc)
            rows = []
            for _ in range(steps):
                d.step()
                rows.append(d.drain())
            alone.append(state(d))
            alone_rows.append(rows)
            d.close()
        runs = [build(ctx, sc) for sc in scenarios()]
        batch_rows = [[] for _ in runs]
        for _ in range(steps):
            ctx.step_batch(runs)
            for i, d in enumerate(runs):
                batch_rows[i].append(d.drain())
        for i, d in enumerate(runs):
            st = state(d)
            assert st["stats"] == alone[i]["stats"], "run %d Poisson stats" % i
            assert alone[i]["stats"][1] > 0 or scenarios()[i][3] is None
            for k, v in alone[i].items():
                if k == "stats":
                    continue
                a, b = np.asarray(v), np.asarray(st[k])
                assert a.shape == b.shape, "run %d %s shape" % (i, k)
                if a.dtype == np.float64:
                    assert np.array_equal(bits(a), bits(b)), "run %d %s differs" % (i, k)
                else:
                    assert np.array_equal(a, b), "run %d %s differs" % (i, k)
            for step, ((ra, la), (rb, lb)) in enumerate(zip(alone_rows[i], batch_rows[i])):
                assert np.array_equal(ra.view(np.uint8), rb.view(np.uint8)), "run %d screen rows, step %d" % (i, step)
                assert np.array_equal(la.view(np.uint8), lb.view(np.uint8)), "run %d lost rows, step %d" % (i, step)
        assert sum(len(r) for rr in batch_rows for r, _ in rr) > 0
        for d in runs:
            d.close()


def test_moved_run_matches_unmoved_run():
    """A run built in one context, moved (spf_run_move) to a second and stepped
    there, then moved to a third and read, has the bits of a run that never
    moved; the old ids are gone from the contexts it left."""
    sc = scenarios()[0]
    steps = 7
    with Context(0) as a, Context(0) as b, Context(0) as c:
        ref = build(a, sc)
        for _ in range(steps):
            ref.step()
        want = state(ref)
        ref.close()
        d = build(a, sc)
        old_id = d.id
        d.move_to(b)
        assert d.ctx is b
        for _ in range(steps):
            d.step()
        d.move_to(c)
        got = state(d)
        with pytest.raises(RuntimeError):
            a.check(a.lib.spf_run_destroy(a.h, old_id))
        for k, v in want.items():
            if k == "stats":
                assert got[k] == v
                continue
            x, y = np.asarray(v), np.asarray(got[k])
            assert np.array_equal(bits(x) if x.dtype == np.float64 else x, bits(y) if y.dtype == np.float64 else y), k
        d.close()
