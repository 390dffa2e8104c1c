"""The high-resolution single run's grid: the 1.6-cell gun at dz = dr = 25 um
(4201 x 1681 lines, 7.06M cells).

Whole steps at that size are compared bit for bit with the CPU restatement
at 1.5M particles (the restatement's time stays within seconds); at the
run's own 16M particles the device is checked by properties the mathematics
fixes: the deposit's charge-flux sums, particle bookkeeping through losses,
and the bounded energy of the fields without sources.
"""
import numpy as np
import pytest

from _support import bits
from oracle.oracle import OracleSim
from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd import workloads as W
from spiffe_amd.device import Context, DeviceRun

pytestmark = pytest.mark.gpu

FIELDS = ["Ez", "Er", "Bphi", "Bphi2", "Ephi", "Br", "Br2", "Bz", "Bz2", "Jz", "Jr", "Jphi", "Psi", "Q"]
SCREENS = [(0.002, 1), (0.10, 1)]


@pytest.fixture(scope="module")
def hires():
    g = W.gun_grid(dz=W.HIRES_DZ)
    return g, W.gun_flags(g)


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def edge_particles(g, n, seed):
    """Particles that leave during the first steps: half drift back through the
    cathode plane z = zmin, half out through the wall at the top of the half cell."""
    rng = np.random.default_rng(seed)
    h = n // 2
    z = np.concatenate([g.zmin + rng.uniform(2e-6, 60e-6, h), g.zmin + rng.uniform(5e-3, 30e-3, n - h)])
    r = np.concatenate([rng.uniform(0, 1e-3, h), rng.uniform(41.80e-3, 41.99e-3, n - h)])
    pz = np.concatenate([-0.5 * np.ones(h), 0.1 * rng.standard_normal(n - h)])
    pr = np.concatenate([0.01 * rng.standard_normal(h), 0.8 * np.ones(n - h)])
    gam = np.sqrt(pz * pz + pr * pr + 1)
    return {"Q": np.full(n, -1e-17), "z": z, "r": r, "pz": pz, "pr": pr, "r_pphi": np.zeros(n),
            "vz": K.c_mks * pz / gam, "vr": K.c_mks * pr / gam, "vphi": np.zeros(n), "r0": r.copy(),
            "z0": z.copy(), "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(), "zmax": z.copy(),
            "zmin": z.copy(), "particleID": np.arange(n, dtype=np.int32) + 100_000_000,
            "generation": np.zeros(n, dtype=np.int16), "status": np.full(n, abi.PART_ACTIVE, dtype=np.uint32)}


def joined(a, b):
    return {k: np.concatenate([a[k], b[k]]) for k in a}


def same(a, b, what):
    if not np.array_equal(bits(a), bits(b)):
        bad = np.argwhere(bits(a) != bits(b))
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d points, first %s: oracle %r device %r" % (what, len(bad), i, a[i], b[i]))


def test_hires_whole_steps_bitwise(ctx, hires):
    """Ten whole steps on the 4201 x 1681-line gun: TM+TE, the Gaussian solenoid,
    space charge, Poisson corrections at steps 0 and 5 (capped at 60
    half-sweeps so the restatement finishes), a screen inside the bunch and
    particles lost through the cathode plane and the wall; fields, particles,
    screen rows and lost rows bit for bit."""
    g, f = hires
    p = joined(W.gun_bunch_gaussian(g, 1_500_000, charge=W.HIRES_CHARGE * 1_500_000 / W.HIRES_PARTICLES,
                                    radius=W.HIRES_RADIUS), edge_particles(g, 20_000, 5))
    n = len(p["z"])
    o = OracleSim(g, f, max_particles=n, has_solenoid=True)
    d = DeviceRun(ctx, g, f, max_particles=n, screens=SCREENS, compat=False)
    for k, v in W.gun_fields(g).items():
        o.set_field(k, v)
        d.upload(k, v)
    o.define_solenoid(**W.GUN_SOLENOID)
    d.define_solenoid(**W.GUN_SOLENOID)
    o.append(p)
    d.append(p)
    pc = dict(W.HIRES_POISSON, maximum_iterations=60)
    for obj in (o, d):
        obj.setup_poisson_correction(**pc)
    n_rows = n_lost = 0
    for step in range(10):
        ro, lo = o.step(screens=SCREENS, compat=False)
        d.step()
        rd, ld = d.drain()
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
        for name in ["r", "pz", "pr", "pphi", "t", "q", "r0", "t0"]:
            same(ro[name], rd[name], "screen column %s, step %d" % (name, step))
        assert np.array_equal(ro["ip"], rd["ip"]) and np.array_equal(ro["screen"], rd["screen"])
        for name in ["t", "rLost", "zLost", "pz", "pr", "q", "particleID", "ip", "directionCode"]:
            assert np.array_equal(lo[name], ld[name]), "lost column %s, step %d" % (name, step)
        n_rows += len(ro)
        n_lost += len(lo)
    assert n_rows > 10_000 and n_lost > 5_000
    for name in FIELDS:
        same(o.field(name), d.download(name), name)
    po, pd = o.particles(), d.particles()
    assert len(po["z"]) == len(pd["z"]) == n - n_lost
    for name in abi.PARTICLE_DOUBLES:
        same(po[name], pd[name], "particle column " + name)
    for name in ("particleID", "generation", "status"):
        assert np.array_equal(po[name], pd[name])
    d.close()


def test_hires_16M_properties(ctx, hires):
    """The run's own 16M particles (plus 100k that leave): the deposit satisfies
    sum J x cell volume = sum Q v for each component; over six whole steps
    (a correction at steps 0 and 5) every particle removed has a lost row, and
    the ones left keep their charges and ids in order."""
    g, f = hires
    p = joined(W.gun_bunch_gaussian(g, W.HIRES_PARTICLES, charge=W.HIRES_CHARGE, radius=W.HIRES_RADIUS),
               edge_particles(g, 100_000, 6))
    n = len(p["z"])
    d = DeviceRun(ctx, g, f, max_particles=n, screens=SCREENS, compat=False)
    for k, v in W.gun_fields(g).items():
        d.upload(k, v)
    d.define_solenoid(**W.GUN_SOLENOID)
    d.append(p)
    # the deposit's charge-flux sums, src/space_charge.c:34-295
    d.zero_current_arrays()
    d.add_particle_currents()
    ir = np.arange(g.nr + 1)
    vc = np.pi * g.dz * g.dr ** 2 * np.where(ir == 0, 0.25, 2.0 * ir)
    vo = np.pi * g.dz * g.dr ** 2 * (2.0 * ir - 1)
    # a particle deposits where its cell row is below the top two rows (Jz, Jphi:
    # r/dr + 1/2 <= nr-2; Jr: r/dr + 1 <= nr-2), src/space_charge.c:92-107, :173-188
    irh = (p["r"] / g.dr + 0.5).astype(np.int64)
    ira = (p["r"] / g.dr + 1.0).astype(np.int64)
    qv = {k: p["Q"] * p[k] for k in ("vz", "vr", "vphi")}
    assert np.count_nonzero(irh > g.nr - 2) > 100   # the wall particles test the cut
    assert np.sum(d.download("Jz") * vc) == pytest.approx(np.sum(qv["vz"][irh <= g.nr - 2]), rel=1e-9)
    assert np.sum(d.download("Jphi") * vc) == pytest.approx(np.sum(qv["vphi"][irh <= g.nr - 2]),
                                                            abs=1e-12 * np.sum(np.abs(qv["vz"])))
    assert np.sum(d.download("Jr")[:, 1:] * vo[1:]) == pytest.approx(np.sum(qv["vr"][ira <= g.nr - 2]), rel=1e-9)
    d.setup_poisson_correction(**W.HIRES_POISSON)
    lost_ids = []
    screen_rows = 0
    for _ in range(6):
        d.step()
        rows, lost = d.drain()
        lost_ids.append(lost["particleID"])
        screen_rows += len(rows)
    lost_ids = np.concatenate(lost_ids)
    left = d.particles()
    assert len(left["z"]) + len(lost_ids) == n
    assert len(np.unique(lost_ids)) == len(lost_ids)
    assert np.array_equal(np.sort(np.concatenate([left["particleID"], lost_ids])), np.sort(p["particleID"]))
    edge = left["particleID"] >= 100_000_000
    assert np.all(left["Q"][~edge] == p["Q"][0]) and np.all(left["Q"][edge] == p["Q"][-1])
    assert np.all(left["status"] & abi.PART_ACTIVE)
    assert len(lost_ids) > 30_000 and screen_rows > 10_000
    last, solves, _ = d.poisson_stats()
    assert solves == 2
    d.close()


def field_energy(d, g):
    """Stored energy of the TM+TE fields, the leapfrog's conserved form
    eps0 E^n.E^n/2 + B^(n-1/2).B^(n+1/2)/(2 mu0) summed over cell volumes
    2 pi r dr dz: after an advance E and the half-step copy B2 = (B^(n-1/2) +
    B^(n+1/2))/2 sit at level n and B at n+1/2, so B^(n-1/2) = 2 B2 - B."""
    r_int = g.dr * np.arange(g.nr + 1)
    r_half = g.dr * (np.arange(g.nr + 1) - 0.5)
    w_int = 2 * np.pi * np.maximum(r_int, 0.25 * g.dr) * g.dr * g.dz
    w_half = 2 * np.pi * np.abs(r_half) * g.dr * g.dz

    def bb(name, sl, w):
        b, b2 = d.download(name)[sl], d.download(name + "2")[sl]
        return np.sum(b * (2 * b2 - b) * w)

    e = 0.5 * K.epsilon_o * (np.sum(d.download("Ez")[1:g.nz, :g.nr - 1] ** 2 * w_int[:g.nr - 1]) +
                             np.sum(d.download("Er")[:g.nz, 1:g.nr] ** 2 * w_half[1:g.nr]) +
                             np.sum(d.download("Ephi")[:g.nz, 1:g.nr] ** 2 * w_int[1:g.nr]))
    b = 0.5 / K.mu_o * (bb("Bphi", np.s_[1:g.nz, 1:g.nr], w_half[1:g.nr]) +
                        bb("Br", np.s_[1:g.nz, :g.nr], w_int[:g.nr]) +
                        bb("Bz", np.s_[:g.nz, 1:g.nr], w_half[1:g.nr]))
    return e + b


def test_hires_field_energy_bounded(ctx, hires):
    """Fields only, no sources, in the closed gun cavity: over 400 FDTD steps
    the stored energy never grows past its value after the first step."""
    g, f = hires
    d = DeviceRun(ctx, g, f)
    for k, v in W.gun_fields(g).items():
        d.upload(k, v)
    d.upload("Ephi", 1e5 * np.outer(np.sin(np.linspace(0, 6, g.nz + 1)), np.sin(np.linspace(0, 4, g.nr + 1))))
    d.advance_field_solutions()
    e1 = field_energy(d, g)
    energies = []
    for k in range(400):
        d.advance_field_solutions()
        if k % 50 == 49:
            energies.append(field_energy(d, g))
    assert e1 > 0
    assert max(energies) < 1.05 * e1, (e1, energies)
    assert min(energies) > 0.2 * e1, (e1, energies)
    d.close()


def test_hires_cell_sort_steps_track_the_reference(ctx, hires):
    """The same ten steps in a cell-sort run (particles re-sorted at steps 0 and
    5, deposition in the library's own fixed order), against the restatement in
    the reference's order: screen-crossing and loss counts equal every step,
    every particle's cell indices equal (matched by particleID), positions and
    momenta within 1e-9 of their largest magnitude, fields within 1e-9
    relative L2, the same particles on the screen and lost with the same
    direction codes."""
    from test_gpu_cellsort import assert_rows_match, assert_same_cells
    g, f = hires
    p = joined(W.gun_bunch_gaussian(g, 1_500_000, charge=W.HIRES_CHARGE * 1_500_000 / W.HIRES_PARTICLES,
                                    radius=W.HIRES_RADIUS), edge_particles(g, 20_000, 5))
    n = len(p["z"])
    o = OracleSim(g, f, max_particles=n, has_solenoid=True)
    d = DeviceRun(ctx, g, f, max_particles=n, screens=SCREENS, compat=False, cell_sort=5)
    for k, v in W.gun_fields(g).items():
        o.set_field(k, v)
        d.upload(k, v)
    o.define_solenoid(**W.GUN_SOLENOID)
    d.define_solenoid(**W.GUN_SOLENOID)
    o.append(p)
    d.append(p)
    pc = dict(W.HIRES_POISSON, maximum_iterations=60)
    for obj in (o, d):
        obj.setup_poisson_correction(**pc)
    rows = [[], [], [], []]
    for step in range(10):
        ro, lo = o.step(screens=SCREENS, compat=False)
        d.step()
        rd, ld = d.drain()
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
        for lst, x in zip(rows, (ro, rd, lo, ld)):
            lst.append(x)
    ro, rd, lo, ld = (np.concatenate(x) for x in rows)
    assert len(ro) > 10_000 and len(lo) > 5_000
    assert_rows_match(ro, rd, lo, ld)
    assert_same_cells(g, o.particles(), d.particles(), 10)
    for name in FIELDS:
        a, b = o.field(name), d.download(name)
        assert np.linalg.norm(b - a) <= 1e-9 * max(np.linalg.norm(a), 1e-300), name
    d.close()


def test_hires_16M_cell_sort_properties(ctx, hires):
    """The run's own 16M particles in a cell-sort run (re-sorted every 20 steps,
    as the run is configured): after the re-sort the particles are ordered by
    half-cell and are the same particles; the deposit's charge-flux sums hold
    (sum J x cell volume = sum Q v per component); over six whole steps every
    particle removed has a lost row and the rest keep their charges."""
    from test_gpu_cellsort import half_cell_keys
    g, f = hires
    p = joined(W.gun_bunch_gaussian(g, W.HIRES_PARTICLES, charge=W.HIRES_CHARGE, radius=W.HIRES_RADIUS),
               edge_particles(g, 100_000, 6))
    n = len(p["z"])
    d = DeviceRun(ctx, g, f, max_particles=n, screens=SCREENS, compat=False, cell_sort=20)
    for k, v in W.gun_fields(g).items():
        d.upload(k, v)
    d.define_solenoid(**W.GUN_SOLENOID)
    d.append(p)
    d.sort_particles()
    q = d.particles()
    assert np.all(np.diff(half_cell_keys(g, q)) >= 0)
    assert np.array_equal(np.sort(q["particleID"]), np.sort(p["particleID"]))
    d.zero_current_arrays()
    d.add_particle_currents()
    ir = np.arange(g.nr + 1)
    vc = np.pi * g.dz * g.dr ** 2 * np.where(ir == 0, 0.25, 2.0 * ir)
    vo = np.pi * g.dz * g.dr ** 2 * (2.0 * ir - 1)
    irh = (p["r"] / g.dr + 0.5).astype(np.int64)
    ira = (p["r"] / g.dr + 1.0).astype(np.int64)
    qv = {k: p["Q"] * p[k] for k in ("vz", "vr", "vphi")}
    assert np.sum(d.download("Jz") * vc) == pytest.approx(np.sum(qv["vz"][irh <= g.nr - 2]), rel=1e-9)
    assert np.sum(d.download("Jphi") * vc) == pytest.approx(np.sum(qv["vphi"][irh <= g.nr - 2]),
                                                            abs=1e-12 * np.sum(np.abs(qv["vz"])))
    assert np.sum(d.download("Jr")[:, 1:] * vo[1:]) == pytest.approx(np.sum(qv["vr"][ira <= g.nr - 2]), rel=1e-9)
    d.setup_poisson_correction(**W.HIRES_POISSON)
    lost_ids = []
    for _ in range(6):
        d.step()
        rows, lost = d.drain()
        lost_ids.append(lost["particleID"])
    lost_ids = np.concatenate(lost_ids)
    left = d.particles()
    assert len(left["z"]) + len(lost_ids) == n
    assert np.array_equal(np.sort(np.concatenate([left["particleID"], lost_ids])), np.sort(p["particleID"]))
    edge = left["particleID"] >= 100_000_000
    assert np.all(left["Q"][~edge] == p["Q"][0]) and np.all(left["Q"][edge] == p["Q"][-1])
    assert len(lost_ids) > 30_000
    d.close()
