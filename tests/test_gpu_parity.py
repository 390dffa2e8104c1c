"""Device path vs the CPU restatement, bit for bit, through the C ABI.

Every test builds the same seeded state in the oracle and in a DeviceRun,
advances both, and compares raw 64-bit patterns of every field and particle
column, plus the screen and lost rows.
"""
import numpy as np
import pytest

from _support import (TMTE, bits, gun_flags, gun_grid, pillbox_flags, pillbox_grid, random_fields,
                      random_particles)
from oracle.oracle import OracleSim
from spiffe_amd import abi
from spiffe_amd.device import Context, DeviceRun

pytestmark = pytest.mark.gpu

TM_NAMES = ["Ez", "Er", "Bphi", "Bphi2"]
TE_NAMES = ["Ephi", "Br", "Br2", "Bz", "Bz2"]
J_NAMES = ["Jz", "Jr", "Jphi"]
PCOLS = abi.PARTICLE_DOUBLES


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def pair(ctx, grid, flags, cap=0, screens=(), compat=True, has_solenoid=False):
    o = OracleSim(grid, flags, max_particles=cap, has_solenoid=has_solenoid)
    d = DeviceRun(ctx, grid, flags, max_particles=cap, screens=screens, compat=compat, has_solenoid=has_solenoid)
    return o, d


def load_fields(o, d, fields):
    for k, v in fields.items():
        o.set_field(k, v)
        d.upload(k, v)


def assert_fields(o, d, names):
    for n in names:
        a, b = o.field(n), d.download(n)
        if not np.array_equal(bits(a), bits(b)):
            bad = np.argwhere(bits(a) != bits(b))
            i = tuple(bad[0])
            raise AssertionError("%s differs at %d points, first %s: oracle %r device %r"
                                 % (n, len(bad), i, a[i], b[i]))


def assert_particles(o, d):
    po, pd = o.particles(), d.particles()
    assert len(po["z"]) == len(pd["z"]), "particle count oracle %d device %d" % (len(po["z"]), len(pd["z"]))
    for n in PCOLS:
        assert np.array_equal(bits(po[n]), bits(pd[n])), "particle column %s differs" % n
    for n in ("particleID", "generation", "status"):
        assert np.array_equal(po[n], pd[n]), "particle column %s differs" % n


def assert_rows(ro, rd, lo, ld):
    assert len(ro) == len(rd), "screen rows: oracle %d device %d" % (len(ro), len(rd))
    for n in ["r", "pz", "pr", "pphi", "t", "q", "r0", "t0", "rmin", "rmax", "zmin", "zmax"]:
        assert np.array_equal(bits(ro[n]), bits(rd[n])), "screen column %s differs" % n
    for n in ["ip", "particleID", "generation", "screen"]:
        assert np.array_equal(ro[n], rd[n]), "screen column %s differs" % n
    assert len(lo) == len(ld), "lost rows: oracle %d device %d" % (len(lo), len(ld))
    for n in ["t", "rLost", "zLost", "pz", "pr", "r0", "z0", "q", "particleID", "ip", "generation",
              "directionCode"]:
        assert np.array_equal(lo[n], ld[n]), "lost column %s differs" % n


def test_fdtd_pillbox_tm010_10k_steps(ctx):
    """Field-only pillbox (closed, 51 x 81 lines), TM010 start, 10 000 steps: the
    field maps within 1e-10 relative L2 of the restatement's (and, since every
    update keeps the reference's operation order, bit for bit)."""
    from scipy.special import j0
    g = pillbox_grid()
    f = pillbox_flags(g)
    o, d = pair(ctx, g, f)
    r = g.dr * np.arange(g.nr + 1)
    Ez = np.tile(j0(2.404825557695773 * r / g.rmax), (g.nz + 1, 1)) * 1e6
    load_fields(o, d, {"Ez": Ez})
    for _ in range(10000):
        o.advance_field_solutions()
        d.advance_field_solutions()
    for n in TM_NAMES:
        a, b = o.field(n), d.download(n)
        assert np.linalg.norm(b - a) <= 1e-10 * np.linalg.norm(a), n
    assert_fields(o, d, TM_NAMES)
    d.close()


@pytest.mark.parametrize("bc", [(1, 1, 1), (0, 0, 1), (0, 1, 1), (1, 0, 1)])
def test_fdtd_gun_tmte_masks_and_cutoffs(ctx, bc):
    """Gun-like cavity with interior metal, random fields and currents, TM+TE, 25 steps."""
    g = gun_grid()
    g.left_bc, g.right_bc, g.upper_bc = bc
    f = gun_flags(g)
    assert np.count_nonzero(f & abi.FL_LAST_ACTIVE_PT) > 0
    o, d = pair(ctx, g, f)
    fl = random_fields(g, 11)
    fl.update({k: v * 1e-3 for k, v in random_fields(g, 12, names=J_NAMES, scale=1e3).items()})
    fl.update(random_fields(g, 13, names=["Bphi2", "Br2", "Bz2"]))
    load_fields(o, d, fl)
    for _ in range(25):
        o.advance_field_solutions()
        d.advance_field_solutions()
    assert_fields(o, d, TM_NAMES + TE_NAMES)
    d.close()


def test_line_currents_and_reset(ctx):
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    o, d = pair(ctx, g, f)
    for (dr_, pos, s, e, J) in [(0, 0, 1, g.nz - 1, 3.5), (1, 20, 1, 30, -2.25), (2, 5, 10, 40, 1.0e-3),
                                (0, 0, 1, g.nz - 1, 0.125)]:
        o.add_line_current(dr_, pos, s, e, J)
        d.add_line_current(dr_, pos, s, e, J)
    assert_fields(o, d, J_NAMES)
    o.multiply_current_arrays(0.37)
    d.multiply_current_arrays(0.37)
    assert_fields(o, d, J_NAMES)
    o.zero_current_arrays()
    d.zero_current_arrays()
    assert_fields(o, d, J_NAMES)
    d.close()


def _edge_particles(g, n, seed):
    """Particles on grid lines, half-lines, the axis and outside the valid index ranges."""
    rng = np.random.default_rng(seed)
    p = random_particles(g, n, seed)
    k = n // 4
    p["z"][:k] = g.zmin + g.dz * rng.integers(0, g.nz, k) * 0.5
    p["r"][:k] = g.dr * rng.integers(0, 2 * g.nr, k) * 0.5
    p["r"][k:k + 10] = 0.0
    p["r"][k + 10:k + 20] = g.rmax * (1 + rng.random(10))
    p["z"][k + 20:k + 30] = g.zmin - g.dz * rng.random(10) * 3
    p["z"][k + 30:k + 40] = g.zmax + g.dz * rng.random(10) * 3
    return p


@pytest.mark.parametrize("n,seed", [(1, 1), (777, 2), (20000, 3)])
def test_deposit_ngp_ordered_sums(ctx, n, seed):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    o, d = pair(ctx, g, f, cap=n)
    p = _edge_particles(g, n, seed) if n > 100 else random_particles(g, n, seed)
    # many particles share cells: sums of many terms in index order
    p["status"][::7] = 0
    o.append(p)
    d.append(p)
    o.add_particle_currents()
    d.add_particle_currents()
    assert_fields(o, d, J_NAMES)
    d.close()


@pytest.mark.parametrize("n,seed", [(300_001, 5), (4096 * 3, 6)])
def test_deposit_concentrated_bunch(ctx, n, seed):
    """Most particles in a few cells: cell segments thousands of terms long that
    span many sort tiles, plus a ragged last tile (or none)."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    o, d = pair(ctx, g, f, cap=n)
    p = random_particles(g, n, seed)
    rng = np.random.default_rng(seed + 100)
    k = (n * 3) // 4
    p["z"][:k] = g.zmin + g.dz * (40 + rng.random(k) * 2.0)
    p["r"][:k] = g.dr * (1 + rng.random(k) * 2.0)
    p["status"][::11] = 0
    o.append(p)
    d.append(p)
    o.add_particle_currents()
    d.add_particle_currents()
    assert_fields(o, d, J_NAMES)
    d.close()


@pytest.mark.parametrize("nz,nr,n", [(5, 5, 3000), (4201, 1681, 200_000)])
def test_deposit_key_widths(ctx, nz, nr, n):
    """Grids whose (component, cell) keys take one radix pass (8-bit keys) and four passes
    (25-bit keys, the 4201 x 1681-line gun), the bench grid's three being covered above."""
    g = gun_grid(nz=nz, nr=nr, modes=TMTE | abi.SPACE_CHARGE)
    f = np.zeros(g.shape, dtype=np.int16)
    o, d = pair(ctx, g, f, cap=n)
    p = random_particles(g, n, 9, margin=0.0)
    rng = np.random.default_rng(10)
    k = n // 3
    p["z"][:k] = g.zmin + g.dz * (0.3 * g.nz + rng.random(k) * 1.5)
    p["r"][:k] = g.dr * rng.random(k) * 1.5
    o.append(p)
    d.append(p)
    o.add_particle_currents()
    d.add_particle_currents()
    assert_fields(o, d, J_NAMES)
    d.close()


def test_deposit_empty(ctx):
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    o, d = pair(ctx, g, gun_flags(g), cap=10)
    d.add_particle_currents()
    o.add_particle_currents()
    assert_fields(o, d, J_NAMES)
    d.close()


@pytest.mark.parametrize("compat", [True, False])
def test_push_single_step_rows_and_compaction(ctx, compat):
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    n = 5000
    screens = [(0.02, 1), (0.06, 1), (0.05, -1)]
    o, d = pair(ctx, g, f, cap=n, screens=screens, compat=compat)
    fl = random_fields(g, 21, scale=3e7)
    fl.update(random_fields(g, 22, names=["Bphi2", "Br2", "Bz2"], scale=3e7))
    load_fields(o, d, fl)
    p = random_particles(g, n, 23, flags=f, pz_scale=3.0)
    p["status"][::11] = 0          # some inactive particles drift and may activate
    p["z"][::13] = g.zmin - 0.3 * g.dz
    p["vz"][::13] = abs(p["vz"][::13])
    o.append(p)
    d.append(p)
    for step in range(3):
        o.f.time_step = step
        o.f.time = step * g.dt + g.start_time
        d.set_step(step)
        ro, lo = o.advance_particles(screens, compat)
        d.advance_particles()
        rd, ld = d.drain()
        assert_rows(ro, rd, lo, ld)
        assert_particles(o, d)
    d.close()


@pytest.mark.parametrize("compat,n", [(True, 60000), (False, 60000), (False, 1_200_000)])
def test_push_mass_loss_compaction(ctx, compat, n):
    """One step where ~40% of the particles leave the box: lost slots on both sides of
    the new np and many 1024-particle chunks of the ordered compaction.  From 2^20
    particle slots up the rows and the compaction go through the multi-workgroup
    pipeline, with screens crossed by many particles as well."""
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    screens = [(g.zmin + 0.5 * (g.zmax - g.zmin), 1), (g.zmin + 0.3 * (g.zmax - g.zmin), -1)] if n > 100000 else []
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n, screens=screens, compat=compat)
    fl = random_fields(g, 61, scale=1e6)
    load_fields(o, d, fl)
    p = random_particles(g, n, 62, flags=f, pz_scale=0.5)
    rng = np.random.default_rng(63)
    out = rng.random(n) < 0.4
    p["z"][out] = g.zmax - 0.05 * g.dz
    p["pz"][out] = 5.0
    gm = np.sqrt(p["pz"] ** 2 + p["pr"] ** 2 + 1)
    p["vz"] = 299792458.0 * p["pz"] / gm
    p["vr"] = 299792458.0 * p["pr"] / gm
    o.append(p)
    d.append(p)
    for step in range(2):
        o.f.time_step = step
        o.f.time = step * g.dt + g.start_time
        d.set_step(step)
        ro, lo = o.advance_particles(screens, compat)
        d.advance_particles()
        rd, ld = d.drain()
        if step == 0:
            assert len(lo) > 0.3 * n
            assert len(ro) > (500 if screens else -1)
        assert_rows(ro, rd, lo, ld)
        assert_particles(o, d)
    d.close()


def test_full_size_gun_steps(ctx):
    """The benchmark workload at full size (1051 x 421 lines, 1M particles, TM+TE, space
    charge, a screen): three whole steps including one Poisson correction, bit for bit."""
    from spiffe_amd import workloads as W
    g = W.gun_grid()
    f = W.gun_flags(g)
    n = 1_000_000
    o, d = pair(ctx, g, f, cap=n, screens=W.GUN_SCREENS, compat=False)
    load_fields(o, d, W.gun_fields(g))
    p = W.gun_bunch(g, n)
    o.append(p)
    d.append(p)
    pc = dict(W.GUN_POISSON, maximum_iterations=400)
    for obj in (o, d):
        obj.setup_poisson_correction(**pc)
    for step in range(3):
        ro, lo = o.step(screens=W.GUN_SCREENS, compat=False)
        d.step()
        rd, ld = d.drain()
        assert_rows(ro, rd, lo, ld)
    assert_fields(o, d, TM_NAMES + TE_NAMES + J_NAMES + ["Psi", "Q"])
    assert_particles(o, d)
    d.close()


def test_full_step_pic_gun_100_steps(ctx):
    """Whole step (reset, deposit, line current, FDTD, push) on the gun geometry, TM+TE."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 4000
    screens = [(0.07, 1)]
    o, d = pair(ctx, g, f, cap=n, screens=screens, compat=False)
    fl = random_fields(g, 31, scale=2e6)
    load_fields(o, d, fl)
    p = random_particles(g, n, 32, flags=f, pz_scale=1.0, q=-1e-14)
    o.append(p)
    d.append(p)
    rows_o, rows_d, lost_o, lost_d = [], [], [], []
    for step in range(100):
        J = 1e-3 * np.sin(0.1 * step)
        ro, lo = o.step(lines=[(0, 0, 1, g.nz - 1, J)], screens=screens, compat=False)
        d.step(lines=[(0, 0, 1, g.nz - 1, J)])
        rd, ld = d.drain()
        rows_o.append(ro), rows_d.append(rd), lost_o.append(lo), lost_d.append(ld)
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
    assert_fields(o, d, TM_NAMES + TE_NAMES + J_NAMES)
    assert_particles(o, d)
    assert_rows(np.concatenate(rows_o), np.concatenate(rows_d), np.concatenate(lost_o), np.concatenate(lost_d))
    d.close()


def test_poisson_correction_step(ctx):
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    o, d = pair(ctx, g, f, cap=n)
    fl = random_fields(g, 41, names=["Ez", "Er", "Bphi"], scale=1e5)
    load_fields(o, d, fl)
    p = random_particles(g, n, 42, flags=f, q=-1e-13)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=2, accuracy=1e-9, maximum_iterations=4000)
    its = []
    for step in range(4):
        o.step(screens=(), compat=True)
        d.step()
        its.append(o.pc.last_iterations)
    assert any(i > 0 for i in its)
    assert_fields(o, d, ["Ez", "Er", "Bphi", "Psi", "Q"] + ["Jz", "Jr"])
    assert_particles(o, d)
    d.close()


def test_sor_parallel_plate_imposed_field(ctx):
    """Imposed-E solve between two plates at 0 and V0 (DC-gun pattern)."""
    nz, nr = 41, 21
    from spiffe_amd.device import Grid
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=0.02, rmax=0.01, dt=1e-13,
             modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    marks = np.zeros(g.shape, dtype=np.int16)
    marks[0, :nr] = abi.FL_IS_SURFACE
    marks[nz - 1, :nr] = abi.FL_IS_SURFACE
    from spiffe_amd.geometry import derive_flags
    f = derive_flags(marks)
    o, d = pair(ctx, g, f)
    V0 = 1e5
    psi = np.zeros(g.shape)
    zz = g.zmin + g.dz * np.arange(nz + 1)
    psi[:nz, nr - 1] = V0 * zz[:nz] / g.zmax
    psi[nz - 1, :nr] = V0
    o.set_field("PsiImposed", psi)
    d.upload("PsiImposed", psi)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=0, accuracy=1e-10, maximum_iterations=3000)
    o.compute_imposed_Efield(True)
    d.compute_imposed_Efield(True)
    assert_fields(o, d, ["PsiImposed", "EzImposed", "ErImposed"])
    d.close()


@pytest.mark.parametrize("limit,acc", [(2, 1e-3), (5, 1e-12), (8, 1e-12), (9, 1e-12), (17, 1e-12), (131, 1e-12),
                                       (3000, 1e-6), (3000, 1e-9)])
def test_sor_solve_counts_and_bits(ctx, limit, acc):
    """solve_poisson_cyl: same iteration count (or -1) and the same Psi bits, whether the stop
    falls inside a batch of half-sweeps, on its last half-sweep, or at the iteration limit."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    g.left_bc = abi.DIRICHLET
    f = gun_flags(g)
    o, d = pair(ctx, g, f)
    rng = np.random.default_rng(limit)
    psi0 = rng.standard_normal(g.shape)
    rhs = rng.standard_normal(g.shape) * 1e3
    for obj in (o, d):
        obj.set_field("Psi", psi0) if obj is o else obj.upload("Psi", psi0)
        obj.set_field("Q", rhs) if obj is o else obj.upload("Q", rhs)
    it_o, _ = o.solve_poisson_cyl("Psi", "Q", acc, limit)
    it_d = d.solve_poisson_cyl(0, acc, limit)
    assert it_o == it_d
    assert_fields(o, d, ["Psi"])
    d.close()


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("limit,acc", [(13, 1e-12), (3000, 1e-9)])
def test_sor_either_tile_shape_gives_the_same_solve(ctx, shape, limit, acc):
    """The context's SOR tile shape forced to either build-time shape (12 half-sweeps
    per launch on 104x24-point tiles, or 8 on 112x32): the restatement's iteration
    count and Psi bits, as with the shape chosen by size."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    g.left_bc = abi.DIRICHLET
    f = gun_flags(g)
    o, d = pair(ctx, g, f)
    rng = np.random.default_rng(limit + 1)
    psi0 = rng.standard_normal(g.shape)
    rhs = rng.standard_normal(g.shape) * 1e3
    o.set_field("Psi", psi0)
    o.set_field("Q", rhs)
    d.upload("Psi", psi0)
    d.upload("Q", rhs)
    ctx.set_sor_shape(shape)
    try:
        it_d = d.solve_poisson_cyl(0, acc, limit)
    finally:
        ctx.set_sor_shape(-1)
    it_o, _ = o.solve_poisson_cyl("Psi", "Q", acc, limit)
    assert it_o == it_d
    assert_fields(o, d, ["Psi"])
    d.close()


@pytest.mark.parametrize("case", ["tiny", "huge", "patch", "zero"])
def test_sor_division_outside_fast_set(ctx, case):
    """Residuals whose magnitudes leave the fast division's range (below 1e-280 or above
    1e280) in all tiles or in a patch of them, and an all-zero solve: same bits and count."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    g.left_bc = abi.DIRICHLET
    f = gun_flags(g)
    o, d = pair(ctx, g, f)
    rng = np.random.default_rng(7)
    psi0 = rng.standard_normal(g.shape)
    rhs = rng.standard_normal(g.shape) * 1e3
    if case == "tiny":
        psi0 *= 1e-292
        rhs *= 1e-295
    elif case == "huge":
        psi0 *= 1e284
        rhs *= 1e281
    elif case == "patch":
        psi0[60:90, 10:30] *= 1e-293
        rhs[60:90, 10:30] = 0.0
    else:
        psi0[:] = 0.0
        rhs[:] = 0.0
    for obj in (o, d):
        obj.set_field("Psi", psi0) if obj is o else obj.upload("Psi", psi0)
        obj.set_field("Q", rhs) if obj is o else obj.upload("Q", rhs)
    it_o, _ = o.solve_poisson_cyl("Psi", "Q", 1e-300, 40)
    it_d = d.solve_poisson_cyl(0, 1e-300, 40)
    assert it_o == it_d
    assert_fields(o, d, ["Psi"])
    d.close()


def test_two_runs_one_context_match_alone(ctx):
    """Runs sharing a context and interleaving their steps give the bits each gives alone."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    res = []
    for mode in ("alone", "mixed"):
        runs = []
        for seed in (51, 52):
            d = DeviceRun(ctx, g, f, max_particles=2000, screens=[(0.07, 1)], compat=False)
            d.upload("Ez", random_fields(g, seed, names=["Ez"])["Ez"])
            d.append(random_particles(g, 2000, seed + 100, flags=f))
            runs.append(d)
        if mode == "alone":
            for d in runs:
                for _ in range(20):
                    d.step()
        else:
            for _ in range(20):
                for d in runs:
                    d.step()
        res.append([(d.download("Ez"), d.download("Jz"), d.particles()["z"]) for d in runs])
        for d in runs:
            d.close()
    for a, b in zip(res[0], res[1]):
        for x, y in zip(a, b):
            assert np.array_equal(bits(x), bits(y))
