"""Reference options and semantics beyond the plain step, device vs the CPU
restatement bit for bit: integrate's z_forces_start and J_filter_multiplier,
the signed zeros of the charge removal's zero-weight terms and of deposition
onto uncleared currents, the relative error-charge threshold, and runs in two
contexts driven from two host threads at once.
"""
import threading

import numpy as np
import pytest

from _support import TMTE, bits, gun_flags, gun_grid, random_fields, random_particles
from oracle.oracle import OracleSim
from spiffe_amd import abi
from spiffe_amd.device import Context, DeviceRun

pytestmark = pytest.mark.gpu

TM_NAMES = ["Ez", "Er", "Bphi", "Bphi2"]
TE_NAMES = ["Ephi", "Br", "Br2", "Bz", "Bz2"]


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def same(a, b, what):
    if not np.array_equal(bits(a), bits(b)):
        bad = np.argwhere(bits(a) != bits(b))
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d points, first %s: oracle %r device %r" % (what, len(bad), i, a[i], b[i]))


def same_particles(o, d):
    po, pd = o.particles(), d.particles()
    assert len(po["z"]) == len(pd["z"])
    for n in abi.PARTICLE_DOUBLES:
        same(po[n], pd[n], "particle column " + n)
    for n in ("particleID", "generation", "status"):
        assert np.array_equal(po[n], pd[n]), n


def test_z_forces_start(ctx):
    """integrate's z_forces_start inside the box: particles behind it drift without
    forces, those crossing it get the fraction of the kick past it
    (src/advance_particles.c:190-200), through spf_step."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 6000
    zfs = 0.05
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n, screens=[(0.07, 1)], compat=False)
    d.set_z_forces_start(zfs)
    for k, v in random_fields(g, 61, scale=2e6).items():
        o.set_field(k, v)
        d.upload(k, v)
    p = random_particles(g, n, 62, flags=f, pz_scale=1.0, q=-1e-14)
    o.append(p)
    d.append(p)
    for _ in range(40):
        o.step(screens=[(0.07, 1)], compat=False, z_forces_start=zfs)
        d.step()
        d.drain()
    for name in TM_NAMES + TE_NAMES:
        same(o.field(name), d.download(name), name)
    same_particles(o, d)
    d.close()


def signed_zero_fields(g, seed):
    """Fields that are all zero with random signs, so Gauss's law gives -0 charge
    in some cells (src/poisson_correction.c:141-182)."""
    rng = np.random.default_rng(seed)
    out = {}
    for name in ("Ez", "Er"):
        out[name] = np.where(rng.random(g.shape) < 0.5, -0.0, 0.0)
    return out


@pytest.mark.parametrize("q_sign", [-1.0, 1.0])
def test_charge_removal_zero_terms(ctx, q_sign):
    """The charge removal also subtracts the zero-weight terms Q*fb*fr, Q*fa*fl,
    Q*fa*fr (src/poisson_correction.c:222-236): for electrons these turn -0 cells
    into +0.  A threshold above the error charge stops the correction after the
    removal, so Q is compared as removed."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n)
    for k, v in signed_zero_fields(g, 71).items():
        o.set_field(k, v)
        d.upload(k, v)
    p = random_particles(g, n, 72, flags=f, q=q_sign * 1e-13)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=1, accuracy=1e-6, maximum_iterations=50,
                                     error_charge_threshold=1.0)
    it_o = o.add_poisson_correction(True)
    it_d = d.add_poisson_correction(True)
    assert it_o == it_d == -2
    qo = o.field("Q")
    n_negzero = int(np.count_nonzero((qo == 0) & np.signbit(qo)))
    assert n_negzero > 0
    same(qo, d.download("Q"), "Q after the charge removal")
    d.close()


def test_poisson_correction_signed_zero_fields(ctx):
    """A whole correction (solve, E from Psi, charge recomputed) from signed-zero fields."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 2000
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n)
    for k, v in signed_zero_fields(g, 73).items():
        o.set_field(k, v)
        d.upload(k, v)
    p = random_particles(g, n, 74, flags=f, q=-1e-13)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=1, accuracy=1e-4, maximum_iterations=300)
    assert o.add_poisson_correction(True) == d.add_poisson_correction(True)
    for name in ("Ez", "Er", "Psi", "Q"):
        same(o.field(name), d.download(name), name)
    d.close()


@pytest.mark.parametrize("factor", [0.5, 2.0])
def test_relative_error_charge_threshold(ctx, factor):
    """error_charge_threshold < 0 compares |Qerror/Qreal| with |threshold|
    (src/poisson_correction.c:259-265).  The threshold is set a factor away from
    the ratio the oracle reports, so one case corrects and the other skips."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    fl = random_fields(g, 81, names=["Ez", "Er", "Bphi"], scale=1e3)
    p = random_particles(g, n, 82, flags=f, q=-1e-13)
    probe = OracleSim(g, f, max_particles=n)
    for k, v in fl.items():
        probe.set_field(k, v)
    probe.append(p)
    probe.setup_poisson_correction(step_interval=1, accuracy=1e-4, maximum_iterations=200)
    probe.add_poisson_correction(True)
    ratio = abs(probe.pc.Qerror / probe.pc.Qreal)
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n)
    for k, v in fl.items():
        o.set_field(k, v)
        d.upload(k, v)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=1, accuracy=1e-4, maximum_iterations=200,
                                     error_charge_threshold=-ratio * factor)
    it_o = o.add_poisson_correction(True)
    it_d = d.add_poisson_correction(True)
    assert it_o == it_d
    assert (it_o == -2) == (factor > 1)
    for name in ("Ez", "Er", "Psi", "Q"):
        same(o.field(name), d.download(name), name)
    d.close()


def _scenario(dev, g, f, seed, steps, out, key):
    """One run stepped in its own context: fields, particles, a Poisson
    correction every 4 steps; its final state goes to out[key]."""
    c = Context(dev)
    try:
        d = DeviceRun(c, g, f, max_particles=4000, screens=[(0.07, 1)], compat=False)
        for k, v in random_fields(g, seed, scale=1e6).items():
            d.upload(k, v)
        d.append(random_particles(g, 4000, seed + 1, flags=f, q=-1e-13))
        d.setup_poisson_correction(step_interval=4, accuracy=1e-5, maximum_iterations=400)
        rows = []
        for _ in range(steps):
            d.step()
            rows.append(d.drain()[0])
        out[key] = {name: d.download(name) for name in ("Ez", "Er", "Bphi", "Jz", "Psi", "Q")}
        out[key]["z"] = d.particles()["z"]
        out[key]["rows"] = np.concatenate(rows)
        d.close()
    finally:
        c.close()


def test_two_contexts_two_threads_match_alone():
    """Two runs, each in its own context, stepped from two host threads at the
    same time (ctypes releases the interpreter lock during library calls), end
    with the bits each gives alone.  Each run does Poisson corrections with the
    charge removal, the phase where concurrency once changed results."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    steps = 24
    alone, mixed = {}, {}
    for seed in (91, 93):
        _scenario(0, g, f, seed, steps, alone, seed)
    errors = []

    def worker(seed):
        try:
            _scenario(0, g, f, seed, steps, mixed, seed)
        except Exception as e:   # surfaced below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(s,)) for s in (91, 93)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for seed in (91, 93):
        for name in ("Ez", "Er", "Bphi", "Jz", "Psi", "Q", "z"):
            same(alone[seed][name], mixed[seed][name], "seed %d %s" % (seed, name))
        assert np.array_equal(alone[seed]["rows"].view(np.uint8), mixed[seed]["rows"].view(np.uint8))


def _serial_abs_drive(rhs, nz, nr):
    """abs_total_driving_term of src/poisson.c:37-41: iz outer, ir inner, from +0."""
    s = 0.0
    for iz in range(nz):
        for v in rhs[iz, :nr - 1]:
            s += abs(float(v))
    return s


def _accuracy_near(target, drive, above):
    """(accuracy a, threshold RN(a*drive)) with the threshold the closest one a
    can produce at or below `target` (above=False) or strictly above it (Python
    floats are IEEE doubles, so a*drive rounds as the reference's product)."""
    a0 = target / drive
    cands = []
    a = a0
    for _ in range(64):
        a = np.nextafter(a, -np.inf)
    for _ in range(129):
        cands.append((a * drive, a))
        a = np.nextafter(a, np.inf)
    if above:
        thr, a = min((t, x) for t, x in cands if t > target)
    else:
        thr, a = max((t, x) for t, x in cands if t <= target)
    return float(a), float(thr)


@pytest.mark.parametrize("edge", ["equal", "above"])
def test_sor_stop_at_the_rounding_edge(ctx, edge):
    """The stopping rule src/poisson.c:125 compares a serial sum of |residual|
    with RN(accuracy * serial sum of |rhs|).  The accuracy is chosen so that the
    threshold is the closest one it can produce at or below the reference's
    residual total at half-sweep n (no stop there: the test is strict) or just
    above it (stop), a few ulps away, where another summation order can decide
    the other way; the device must return the reference's iteration count and
    Psi bits."""
    g = gun_grid(modes=abi.TM_FIELDS)
    f = gun_flags(g)
    rng = np.random.default_rng(101)
    rhs = rng.standard_normal(g.shape) * 1e3
    # the reference's residual total after each half-sweep k (a solve limited
    # to k+1 half-sweeps reports it), and a half-sweep n below every earlier
    # total with a smaller one after it, so the threshold decides n alone
    hist = []
    for limit in range(2, 90):
        probe = OracleSim(g, f)
        probe.set_field("Q", rhs)
        it, res = probe.solve_poisson_cyl("Psi", "Q", 1e-300, limit)
        assert it == -1
        hist.append(res)   # residual of half-sweep limit-1
    n = next(k for k in range(20, len(hist)) if hist[k - 1] < min(hist[:k - 1]) and hist[k] < hist[k - 1])
    res_n = hist[n - 1]
    drive = _serial_abs_drive(rhs, g.nz, g.nr)
    acc, thr = _accuracy_near(res_n, drive, edge == "above")
    assert abs(thr - res_n) < 1e-14 * res_n
    o = OracleSim(g, f)
    d = DeviceRun(ctx, g, f)
    o.set_field("Q", rhs)
    d.upload("Q", rhs)
    it_o, _ = o.solve_poisson_cyl("Psi", "Q", acc, 200)
    it_d = d.solve_poisson_cyl(0, acc, 200)
    assert it_o == (n if edge == "above" else n + 1)
    assert it_d == it_o
    same(o.field("Psi"), d.download("Psi"), "Psi")
    d.close()


@pytest.mark.parametrize("mode", ["absolute", "absolute_above", "relative", "relative_above"])
def test_error_charge_threshold_at_the_rounding_edge(ctx, mode):
    """Thresholds equal to the reference's |Qerror| (or |Qerror/Qreal|), or the
    next double above: the strict test skips the correction only in the second
    case (src/poisson_correction.c:259-265), and the device's decision must agree."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    fl = random_fields(g, 111, names=["Ez", "Er", "Bphi"], scale=1e3)
    p = random_particles(g, n, 112, flags=f, q=-1e-13)
    probe = OracleSim(g, f, max_particles=n)
    for k, v in fl.items():
        probe.set_field(k, v)
    probe.append(p)
    probe.setup_poisson_correction(step_interval=1, accuracy=1e-4, maximum_iterations=100)
    probe.add_poisson_correction(True)
    if mode.startswith("absolute"):
        thr = abs(probe.pc.Qerror)
    else:
        thr = abs(probe.pc.Qerror / probe.pc.Qreal)
    if mode.endswith("above"):
        thr = float(np.nextafter(thr, np.inf))
    if mode.startswith("relative"):
        thr = -thr
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n)
    for k, v in fl.items():
        o.set_field(k, v)
        d.upload(k, v)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=1, accuracy=1e-4, maximum_iterations=100,
                                     error_charge_threshold=thr)
    it_o = o.add_poisson_correction(True)
    it_d = d.add_poisson_correction(True)
    assert (it_o == -2) == mode.endswith("above")
    assert it_o == it_d
    for name in ("Ez", "Er", "Psi", "Q"):
        same(o.field(name), d.download(name), name)
    d.close()


@pytest.mark.parametrize("batched", [False, True])
def test_j_filter_steps(ctx, batched):
    """integrate's J_filter_multiplier (src/integrate.c:110-116, :168-178): J is
    scaled by m/(1-m) instead of cleared, deposited on, and scaled by 1-m;
    whole steps (with line currents when stepped alone, through spf_step_batch
    otherwise) match the oracle bit for bit."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    m = 0.3
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n, compat=False)
    for k, v in random_fields(g, 61).items():
        o.set_field(k, v)
        d.upload(k, v)
    p = random_particles(g, n, 62, flags=f)
    o.append(p)
    d.append(p)
    d.set_j_filter(m)
    lines = () if batched else ((0, 3, 10, 40, 2.5e3), (1, 20, 5, 9, -1.5e3))
    for _ in range(8):
        o.step(lines=lines, screens=(), compat=False, j_filter=m)
        if batched:
            ctx.step_batch([d])
        else:
            d.step(lines=lines)
    for name in TM_NAMES + TE_NAMES + ["Jz", "Jr", "Jphi"]:
        same(o.field(name), d.download(name), name)
    same_particles(o, d)
    d.close()


def test_deposit_onto_negative_zero_currents(ctx):
    """A deposition onto currents that hold -0 (not cleared first, as with the J
    filter): the reference's nearest-point loops add +0 at points below the top
    rows and on Jr's row 0, turning those -0 into +0; the device gives the same
    signed zeros."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 4000
    o = OracleSim(g, f, max_particles=n)
    d = DeviceRun(ctx, g, f, max_particles=n)
    for name in ("Jz", "Jr", "Jphi"):
        o.set_field(name, np.full(g.shape, -0.0))
        d.upload(name, np.full(g.shape, -0.0))
    p = random_particles(g, n, 63)
    rng = np.random.default_rng(64)
    k = n // 2
    p["r"][:k] = g.dr * (g.nr - 1.5 + rng.random(k) * 1.4)   # rows nr-1 and nr
    p["r"][k:k + 50] = -0.4 * g.dr * rng.random(50)           # Jr's row 0
    o.append(p)
    d.append(p)
    o.add_particle_currents()
    d.add_particle_currents()
    for name in ("Jz", "Jr", "Jphi"):
        a = o.field(name)
        assert np.any(bits(a) == bits(np.float64(-0.0))) and np.any(bits(a) == bits(np.float64(0.0)))
        same(a, d.download(name), name)
    d.close()
