"""Cell-sort runs (spf_run_desc.cell_sort > 0) against the CPU restatement.

A cell-sort run keeps its particles sorted by cell (re-sorted every N steps)
and deposits in a fixed order of its own: each wavefront's contributions to a
cell are added in one fixed tree and the partials are folded in storage
order.  Its sums are therefore the reference's sums in another order, and the
checks are those the design targets state for floating-point results, with
the tolerance written in each test:
- deposition: per cell, |J - J_ref| <= 2 gamma_(n-1) * sum|contribution|,
  the bound between any two orders of the same n additions (u = 2^-53,
  gamma_k = k u / (1 - k u));
- whole steps: particle, screen-crossing and loss counts equal step by step,
  every particle's cell indices equal (particles matched by particleID),
  fields within 1e-9 relative L2, bunch moments at the screen within 1e-6;
- determinism: the same run twice, and a run stepped in a batch, give the
  same bits.
"""
import math

import numpy as np
import pytest

from _support import TMTE, bits, gun_flags, gun_grid, pillbox_flags, pillbox_grid, random_fields, random_particles
from oracle.oracle import OracleSim
from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd.device import DeviceRun, SpfError
from spiffe_amd.moments import relative_differences, screen_moments
from test_gpu_parity import ctx, load_fields  # noqa: F401

pytestmark = pytest.mark.gpu

U = 2.0 ** -53
J_NAMES = ["Jz", "Jr", "Jphi"]


def gamma(n):
    return n * U / (1 - n * U)


def by_id(p):
    o = np.argsort(p["particleID"], kind="stable")
    return {k: v[o] for k, v in p.items()}


def half_cell_keys(g, p):
    hz = np.clip((2.0 * ((p["z"] - g.zmin) / g.dz)).astype(np.int64), 0, 2 * g.nz + 1)
    hr = np.clip((2.0 * (p["r"] / g.dr)).astype(np.int64), 0, 2 * g.nr + 1)
    return hr * (2 * g.nz + 2) + hz


def abs_bunch(p):
    q = dict(p)
    for k in ("Q", "vz", "vr", "vphi"):
        q[k] = np.abs(p[k])
    return q


def test_cell_sort_needs_compat_off(ctx):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    with pytest.raises(SpfError) as e:
        DeviceRun(ctx, g, gun_flags(g), max_particles=10, compat=True, cell_sort=5)
    assert e.value.code == abi.SPF_EINVAL
    d = DeviceRun(ctx, g, gun_flags(g), max_particles=10, compat=False)
    with pytest.raises(SpfError):
        d.sort_particles()
    d.close()


@pytest.mark.parametrize("n,seed", [(1, 1), (50_000, 2), (300_001, 3)])
def test_sort_keeps_every_particle(ctx, n, seed):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    d = DeviceRun(ctx, g, f, max_particles=n, compat=False, cell_sort=7)
    p = random_particles(g, n, seed, flags=f)
    p["status"][::5] = 0
    d.append(p)
    d.sort_particles()
    q = d.particles()
    keys = half_cell_keys(g, q)
    assert np.all(np.diff(keys) >= 0), "storage is not ordered by half-cell"
    a, b = by_id(p), by_id(q)
    for k in abi.PARTICLE_DOUBLES:
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    for k in ("particleID", "generation", "status"):
        assert np.array_equal(a[k], b[k]), k
    # the sort is stable: within a half-cell, the earlier order is kept
    for k0 in np.unique(keys)[:50]:
        ids = q["particleID"][keys == k0]
        assert np.all(np.diff(ids) > 0)
    d.close()


@pytest.mark.parametrize("n,seed,bunch", [(777, 4, False), (20_000, 5, False), (300_001, 6, True)])
def test_reduce_deposit_within_reordering_bound(ctx, n, seed, bunch):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    p = random_particles(g, n, seed)
    if bunch:   # most particles in a few cells: long per-cell sums
        rng = np.random.default_rng(seed + 100)
        k = (n * 3) // 4
        p["z"][:k] = g.zmin + g.dz * (40 + rng.random(k) * 2.0)
        p["r"][:k] = g.dr * (1 + rng.random(k) * 2.0)
    p["status"][::11] = 0
    o = OracleSim(g, f, max_particles=n)
    o.append(p)
    o.add_particle_currents()
    oa = OracleSim(g, f, max_particles=n)
    oa.append(abs_bunch(p))
    oa.add_particle_currents()
    d = DeviceRun(ctx, g, f, max_particles=n, compat=False, cell_sort=3)
    d.append(p)
    d.sort_particles()
    d.add_particle_currents()
    first = {k: d.download(k) for k in J_NAMES}
    for k in J_NAMES:
        ref, mag = o.field(k), oa.field(k)
        bound = 2 * gamma(n) * mag
        err = np.abs(first[k] - ref)
        assert np.all(err <= bound), "%s: worst excess %g" % (k, np.max(err - bound))
        assert np.linalg.norm(first[k] - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300)
    # the order is fixed: the same deposit again gives the same bits
    d.zero_current_arrays()
    d.add_particle_currents()
    for k in J_NAMES:
        assert np.array_equal(bits(first[k]), bits(d.download(k))), k
    d.close()


def step_pair(ctx, g, f, p, fields, steps, screens, sort_every, pc=None, sol=None, check_every=1):
    """The oracle (reference order) and a cell-sort device run stepped side by side;
    returns both, the per-step event counts and the rows."""
    cap = len(p["z"])
    o = OracleSim(g, f, max_particles=cap, has_solenoid=sol is not None)
    d = DeviceRun(ctx, g, f, max_particles=cap, screens=screens, compat=False, has_solenoid=sol is not None,
                  cell_sort=sort_every)
    for k, v in fields.items():
        o.set_field(k, v)
        d.upload(k, v)
    if sol is not None:
        o.define_solenoid(**sol)
        d.define_solenoid(**sol)
    o.append(p)
    d.append(p)
    if pc is not None:
        o.setup_poisson_correction(**pc)
        d.setup_poisson_correction(**pc)
    ro_all, rd_all, lo_all, ld_all = [], [], [], []
    for step in range(steps):
        ro, lo = o.step(screens=screens, compat=False)
        d.step()
        rd, ld = d.drain()
        assert len(ro) == len(rd), "step %d: screen rows oracle %d device %d" % (step, len(ro), len(rd))
        assert len(lo) == len(ld), "step %d: lost rows oracle %d device %d" % (step, len(lo), len(ld))
        assert d.count()[0] == len(o.particles()["z"]), "step %d: particle count" % step
        if step % check_every == 0:
            assert_same_cells(g, o.particles(), d.particles(), step)
        ro_all.append(ro), rd_all.append(rd), lo_all.append(lo), ld_all.append(ld)
    return o, d, np.concatenate(ro_all), np.concatenate(rd_all), np.concatenate(lo_all), np.concatenate(ld_all)


def assert_same_cells(g, po, pd, step, rel=1e-9):
    a, b = by_id(po), by_id(pd)
    assert np.array_equal(a["particleID"], b["particleID"]), "step %d: different particles" % step
    for name, x, h in (("iz", "z", g.dz), ("ir", "r", g.dr)):
        ca = ((a[x] - (g.zmin if x == "z" else 0.0)) / h).astype(np.int64)
        cb = ((b[x] - (g.zmin if x == "z" else 0.0)) / h).astype(np.int64)
        bad = np.flatnonzero(ca != cb)
        assert len(bad) == 0, "step %d: %s differs for %d particles (first id %d)" % (
            step, name, len(bad), a["particleID"][bad[0]])
    for k in ("z", "r", "pz", "pr"):
        scale = np.max(np.abs(a[k])) or 1.0
        assert np.max(np.abs(a[k] - b[k])) <= rel * scale, "step %d: %s beyond %g" % (step, k, rel)
    assert np.array_equal(a["status"], b["status"])


def assert_rows_match(ro, rd, lo, ld):
    assert np.array_equal(np.sort(ro["particleID"]), np.sort(rd["particleID"]))
    assert np.array_equal(np.sort(lo["particleID"]), np.sort(ld["particleID"]))
    if len(lo):
        a = lo[np.argsort(lo["particleID"], kind="stable")]
        b = ld[np.argsort(ld["particleID"], kind="stable")]
        assert np.array_equal(a["directionCode"], b["directionCode"])
    if len(ro) > 20:
        diffs = relative_differences(screen_moments(ro), screen_moments(rd))
        assert max(diffs.values()) <= 1e-6, diffs


def test_cell_sort_gun_steps_track_the_reference(ctx):
    """Gun cavity, TM+TE, space charge, Poisson correction every 5 steps, a
    solenoid, a screen the bunch crosses and particles driven into the walls;
    40 steps with a re-sort every 4."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 30_000
    p = random_particles(g, n, 21, flags=f, pz_scale=2.0)
    rng = np.random.default_rng(22)
    k = n // 2   # a dense bunch near the screen, moving forward
    p["z"][:k] = g.zmin + g.dz * (60 + rng.random(k) * 3.0)
    p["r"][:k] = g.dr * (0.5 + rng.random(k) * 4.0)
    p["pz"][:k] = 3.0 + 0.1 * rng.standard_normal(k)
    gm = np.sqrt(1 + p["pz"] ** 2 + p["pr"] ** 2)
    p["vz"] = K.c_mks * p["pz"] / gm
    p["vr"] = K.c_mks * p["pr"] / gm
    fields = random_fields(g, 23, scale=3e5)
    screens = [(g.zmin + g.dz * 64.5, 1)]
    pc = dict(step_interval=5, accuracy=1e-6, maximum_iterations=400)
    sol = dict(gaussian=1, Bz_peak=0.2, gaussian_center=0.07, gaussian_sigma=0.02)
    o, d, ro, rd, lo, ld = step_pair(ctx, g, f, p, fields, 40, screens, 4, pc=pc, sol=sol, check_every=5)
    assert len(ro) > 1000 and len(lo) > 100, (len(ro), len(lo))
    assert_same_cells(g, o.particles(), d.particles(), 40)
    assert_rows_match(ro, rd, lo, ld)
    for name in ("Ez", "Er", "Bphi", "Ephi", "Br", "Bz", "Jz", "Jr", "Jphi", "Psi"):
        a, b = o.field(name), d.download(name)
        assert np.linalg.norm(b - a) <= 1e-9 * max(np.linalg.norm(a), 1e-300), name
    d.close()


def test_cell_sort_pillbox_emission_1000_steps(ctx):
    """The plain small pattern (TM010 pillbox, 1 mm cathode emitting a Gaussian
    pulse of ~10 k macroparticles, space charge, a screen) in a cell-sort run:
    emitted, lost and screen-crossing counts equal every step and cell indices
    equal every 50 steps over 1000 steps, moments at the screen within 1e-6."""
    from scipy.special import j0
    TMI = abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP
    g = pillbox_grid(modes=TMI | abi.SPACE_CHARGE)
    f = pillbox_flags(g)
    screens = [(0.005, 1)]
    o = OracleSim(g, f, max_particles=16000)
    d = DeviceRun(ctx, g, f, max_particles=16000, screens=screens, compat=False, cell_sort=20)
    r = g.dr * np.arange(g.nr + 1)
    Ez = np.tile(-2.0e7 * j0(2.404825557695773 * r / g.rmax), (g.nz + 1, 1))
    Ez[:, g.nr - 1:] = 0
    load_fields(o, d, {"Ez": Ez})
    sigma, Q = 2e-12, 0.1e-9
    epm = Q / K.e_mks / 10000
    J = Q / (sigma * math.sqrt(2 * math.pi)) / (math.pi * 1e-6)
    kw = dict(z_position=0.0, outer_radius=1e-3, current_density=J, electrons_per_macroparticle=epm,
              initial_pz=1.978e-3, zoned_emission=0, random_number_seed=987654321)
    t = 4 * sigma + sigma * np.linspace(-4, 4, 81)
    prof = (t, np.exp(-0.5 * ((t - 4 * sigma) / sigma) ** 2))
    o.define_cathode(profile=prof, **kw)
    d.define_cathode(profile=prof, **kw)
    rows = [[], [], [], []]
    for step in range(1000):
        ro, lo = o.step(screens=screens, compat=False)
        d.step()
        rd, ld = d.drain()
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
        n = d.count()[0]
        assert n == len(o.particles()["z"]), "particle count differs at step %d" % step
        if step % 50 == 0 and n:
            assert_same_cells(g, o.particles(), d.particles(), step)
        for lst, x in zip(rows, (ro, rd, lo, ld)):
            lst.append(x)
    ro, rd, lo, ld = (np.concatenate(x) for x in rows)
    assert len(rd) > 1000
    assert_rows_match(ro, rd, lo, ld)
    d.close()


def test_cell_sort_run_is_deterministic_alone_and_batched(ctx):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    p = random_particles(g, 40_000, 31, flags=f)
    fields = random_fields(g, 32, scale=3e5)

    def make():
        d = DeviceRun(ctx, g, f, max_particles=40_000, compat=False, cell_sort=3)
        for k, v in fields.items():
            d.upload(k, v)
        d.append(p)
        return d

    a, b, c = make(), make(), make()
    other = DeviceRun(ctx, g, f, max_particles=1000, compat=False)
    other.append(random_particles(g, 1000, 33, flags=f))
    for _ in range(10):
        a.step()
        b.step()
        ctx.step_batch([c, other])
    for name in ("Jz", "Jr", "Jphi", "Ez", "Bphi"):
        x = bits(a.download(name))
        assert np.array_equal(x, bits(b.download(name))), name
        assert np.array_equal(x, bits(c.download(name))), name
    pa, pc_ = a.particles(), c.particles()
    for k in ("z", "r", "pz", "pr", "particleID"):
        assert np.array_equal(pa[k].view(np.uint8), pc_[k].view(np.uint8)), k
    for r in (a, b, c, other):
        r.close()


def test_cell_sort_runs_deposit_together_with_their_own_bits(ctx):
    """Three cell-sort runs on different grids (and an exact-order run) stepped
    together: the cell-sort runs' deposition goes through one set of launches
    over a shared key space; every run must keep the bits it has alone."""
    from _support import gun_grid as gg
    specs = [(141, 61, 30_000, 41), (101, 45, 12_000, 42), (181, 77, 50_000, 43)]

    def make(nz, nr, n, seed, cs=4):
        g = gg(nz=nz, nr=nr, modes=TMTE | abi.SPACE_CHARGE)
        f = gun_flags(g)
        d = DeviceRun(ctx, g, f, max_particles=n, compat=False, cell_sort=cs, screens=[(0.07, 1)])
        for k, v in random_fields(g, seed, scale=3e5).items():
            d.upload(k, v)
        d.append(random_particles(g, n, seed + 1, flags=f, pz_scale=1.0))
        return d

    alone = [make(*sp) for sp in specs]
    batch = [make(*sp) for sp in specs]
    exact = make(121, 51, 8_000, 44, cs=0)
    for _ in range(12):
        for d in alone:
            d.step()
        ctx.step_batch(batch + [exact])
    for a, b in zip(alone, batch):
        for name in ("Jz", "Jr", "Jphi", "Ez", "Er", "Bphi"):
            assert np.array_equal(bits(a.download(name)), bits(b.download(name))), name
        pa, pb = a.particles(), b.particles()
        for k in ("z", "r", "pz", "pr", "r_pphi"):
            assert np.array_equal(bits(pa[k]), bits(pb[k])), k
        assert np.array_equal(pa["particleID"], pb["particleID"])
    for d in alone + batch + [exact]:
        d.close()
