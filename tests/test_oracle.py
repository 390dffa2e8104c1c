"""Pins the CPU restatement (the checker) with closed forms, invariants and
brute force on tiny inputs.  The reference ships no tests or golden vectors,
so each check is chosen so that a plausible slip in the restatement (a wrong
sign, index, stagger offset or coefficient) fails it."""
import math

import numpy as np
import pytest
from scipy.special import j0, j1

from _support import TMTE, gun_flags, gun_grid, pillbox_flags, pillbox_grid, random_particles
from oracle.oracle import OracleSim, derive_flags as c_derive_flags
from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd.device import Grid, courant_dt
from spiffe_amd.geometry import derive_flags, pillbox_marks

J0_1 = 2.404825557695773   # first zero of J0
J1_1 = 3.831705970207512   # first zero of J1


# ---------------------------------------------------------------- flags
def test_flag_derivation_hand_case():
    """Hand-derived flags of src/geometry.c:319-362 on a 7 x 5 lattice."""
    nz, nr = 6, 4
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[1, 0] = m[2, 0] = m[2, 1] = abi.FL_IS_SURFACE
    m[3, 3] = abi.FL_IS_SURFACE
    m[4, 3] = m[5, 3] = abi.FL_IS_INSIDE
    m[0:6, 2] = abi.FL_IS_INSIDE          # an all-metal row
    m[1:6, 1] |= abi.FL_IS_INSIDE         # a row metal from iz=1 to the end
    f = c_derive_flags(m)
    E, EZ, ER, BZ, BR, BP, L = (abi.FL_EPHI_ZERO, abi.FL_EZ_ZERO, abi.FL_ER_ZERO, abi.FL_BZ_ZERO,
                                abi.FL_BR_ZERO, abi.FL_BPHI_ZERO, abi.FL_LAST_ACTIVE_PT)
    # (1,0): metal, nothing metal below in z, on axis -> Ephi, Er, Bz zero
    assert f[1, 0] == abi.FL_IS_SURFACE | E | ER | BZ
    # (2,0): iz-1 metal -> Ez, Br; axis -> Er, Bz; iz-1 metal and axis -> Bphi
    assert f[2, 0] == abi.FL_IS_SURFACE | E | EZ | BR | ER | BZ | BP
    # (3,3): nothing metal at iz-1 or ir-1 ((3,2) is inside metal!) -> Er, Bz from the r neighbour
    assert f[3, 3] == abi.FL_IS_SURFACE | E | ER | BZ
    # (4,3): (3,3) metal, (4,2) metal, (3,2) metal -> everything incl. Bphi
    assert f[4, 3] & (EZ | BR | ER | BZ | BP) == EZ | BR | ER | BZ | BP
    # last-active points: row ir=2 is all metal -> flag at iz=0; row ir=3 ends metal from iz=4 -> flag at 4;
    # row ir=1 is metal from iz=1 -> the scan stops at iz=0 and no flag is set
    lap = np.argwhere(f & L)
    assert sorted(map(tuple, lap)) == [(0, 2), (4, 3)]


def test_flag_derivation_two_implementations_agree():
    """The product's numpy derivation and the C restatement agree on random masks."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        nz, nr = rng.integers(3, 40, 2)
        m = rng.choice([0, 0, 0, abi.FL_IS_SURFACE, abi.FL_IS_INSIDE], size=(nz + 1, nr + 1)).astype(np.int16)
        assert np.array_equal(derive_flags(m), c_derive_flags(m))


# ---------------------------------------------------------------- FDTD
def walled_pillbox(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP, R=0.04, L=0.025, d=0.5e-3, extra=6):
    """Pillbox whose radial wall is metal inside the grid (r >= R), so the result does not depend
    on the upper-boundary patch of src/integrate.c:406, which acts as a wall one cell below rmax."""
    nz = int(round(L / d)) + 1
    nr = int(round(R / d)) + 1 + extra
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=L, rmax=d * (nr - 1), dt=courant_dt(d, d, 0.9), modes=modes)
    m = np.zeros(g.shape, dtype=np.int16)
    iw = int(round(R / d))
    m[:nz, iw] = abi.FL_IS_SURFACE
    m[:nz, iw + 1:nr] = abi.FL_IS_INSIDE
    return g, derive_flags(m), R


def _ring_frequency(series, dt):
    """Frequency of the dominant oscillation from zero crossings of a sampled signal."""
    s = np.asarray(series)
    idx = np.nonzero(np.sign(s[:-1]) != np.sign(s[1:]))[0]
    t = (idx + s[idx] / (s[idx] - s[idx + 1])) * dt   # linear interpolation of the crossing
    return 1.0 / (2 * np.mean(np.diff(t)))


def test_tm010_pillbox_frequency():
    """TM010 of a closed pillbox: f = 2.405 c / (2 pi R) (2.8687 GHz for R = 4 cm)."""
    g, f, R = walled_pillbox()
    o = OracleSim(g, f)
    r = np.minimum(g.dr * np.arange(g.nr + 1), R)
    o.set_field("Ez", np.tile(j0(J0_1 * r / R), (g.nz + 1, 1)))
    probe = []
    for _ in range(1500):
        o.advance_field_solutions()
        probe.append(o.arr["Ez"][g.nz // 2, 3])
    fr = _ring_frequency(probe, g.dt)
    f0 = J0_1 * K.c_mks / (2 * np.pi * R)
    assert abs(fr / f0 - 1) < 1e-3, (fr, f0)
    # Er stays zero for a z-independent TM0n0 mode; Bphi is not
    assert np.abs(o.arr["Er"]).max() < 1e-9 * np.abs(o.arr["Ez"]).max()
    assert np.abs(o.arr["Bphi"]).max() > 0


def test_tm011_pillbox_frequency_exercises_er():
    """TM011: f = c/(2 pi) sqrt((2.405/R)^2 + (pi/L)^2); needs the Er update."""
    g, fl, R = walled_pillbox()
    o = OracleSim(g, fl)
    L = g.zmax - g.zmin
    r = np.minimum(g.dr * np.arange(g.nr + 1), R)
    z_ez = g.zmin + g.dz * (np.arange(g.nz + 1) - 0.5)   # Ez stagger, src/spiffe.h:60
    z_er = g.zmin + g.dz * np.arange(g.nz + 1)           # Er stagger, src/spiffe.h:61
    r_er = np.clip(g.dr * (np.arange(g.nr + 1) - 0.5), 0, R)
    kz, kr = np.pi / L, J0_1 / R
    # divergence-free TM011 electric field: Er = (kz/kr) J1(kr r) sin(kz z)
    o.set_field("Ez", np.outer(np.cos(kz * z_ez), j0(kr * r)))
    o.set_field("Er", (kz / kr) * np.outer(np.sin(kz * z_er), j1(kr * r_er)))
    probe = []
    for _ in range(1500):
        o.advance_field_solutions()
        probe.append(o.arr["Ez"][3, 2])
    fr = _ring_frequency(probe, g.dt)
    f0 = K.c_mks / (2 * np.pi) * np.hypot(J0_1 / R, np.pi / L)
    assert abs(fr / f0 - 1) < 2e-3, (fr, f0)
    assert np.abs(o.arr["Er"]).max() > 0.1 * np.abs(o.arr["Ez"]).max()


def _te011_error(d):
    g, fl, R = walled_pillbox(modes=abi.TE_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP, d=d)
    o = OracleSim(g, fl)
    L = g.zmax - g.zmin
    r = np.minimum(g.dr * np.arange(g.nr + 1), R)
    z = g.zmin + g.dz * np.arange(g.nz + 1)              # Ephi stagger, src/spiffe.h:67
    o.set_field("Ephi", np.outer(np.sin(np.pi * z / L), j1(J1_1 * r / R)))
    probe = []
    for _ in range(int(round(1200 * 0.5e-3 / d))):
        o.advance_field_solutions()
        probe.append(o.arr["Ephi"][g.nz // 2, int(round(0.02 / d))])
    fr = _ring_frequency(probe, g.dt)
    return fr / (K.c_mks / (2 * np.pi) * np.hypot(J1_1 / R, np.pi / L)) - 1


def test_te011_pillbox_frequency_first_order():
    """TE011: f -> c/(2 pi) sqrt((3.832/R)^2 + (pi/L)^2).  The reference's Bz update
    (src/integrate.c:515-520) weights Ephi as if it sat one cell further out than
    src/spiffe.h:67 places it, a first-order error: it must halve when dr halves."""
    e1, e2 = _te011_error(1e-3), _te011_error(0.5e-3)
    assert abs(e2) < 6e-3
    assert 1.8 < e1 / e2 < 2.2, (e1, e2)


def _field_energy(o, g):
    r_i = g.dr * np.arange(g.nr + 1)
    r_h = g.dr * (np.arange(g.nr + 1) - 0.5)
    w_i = 2 * np.pi * np.maximum(r_i, g.dr / 4) * g.dr * g.dz
    w_h = 2 * np.pi * np.abs(r_h) * g.dr * g.dz
    a = o.arr
    return (0.5 * K.epsilon_o * (np.sum(a["Ez"][1:g.nz, :g.nr - 1] ** 2 * w_i[:g.nr - 1]) +
                                 np.sum(a["Er"][:g.nz, 1:g.nr] ** 2 * w_h[1:g.nr])) +
            0.5 / K.mu_o * np.sum(a["Bphi2"][1:g.nz, 1:g.nr] ** 2 * w_h[1:g.nr]))


def test_closed_cavity_energy_bounded():
    """Lossless closed cavity, no sources: stored energy neither grows nor decays."""
    g, fl, R = walled_pillbox()
    o = OracleSim(g, fl)
    r = np.minimum(g.dr * np.arange(g.nr + 1), R)
    L = g.zmax - g.zmin
    z_ez = g.zmin + g.dz * (np.arange(g.nz + 1) - 0.5)
    o.set_field("Ez", np.outer(np.cos(np.pi * z_ez / L), j0(J0_1 * r / R)) +
                np.tile(j0(J0_1 * r / R), (g.nz + 1, 1)))
    energies = []
    for k in range(3000):
        o.advance_field_solutions()
        if k > 200:
            energies.append(_field_energy(o, g))
    e = np.array(energies)
    # energy at the E and B2 time levels oscillates slightly; its envelope must be flat
    first, last = e[:500].max(), e[-500:].max()
    assert abs(last / first - 1) < 0.02
    assert e.min() > 0.5 * e.max()


def test_masked_points_never_change_and_axis_mirror():
    """Points flagged *_ZERO keep their value; axis rows obey Er[0] = -Er[1], Bphi[0] = -Bphi[1]."""
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    o = OracleSim(g, f)
    rng = np.random.default_rng(1)
    init = {n: rng.standard_normal(g.shape) for n in ("Ez", "Er", "Bphi", "Ephi", "Br", "Bz")}
    for n, v in init.items():
        o.set_field(n, v)
    for _ in range(5):
        o.advance_field_solutions()
    a = o.arr
    interior = np.zeros(g.shape, dtype=bool)
    interior[2:g.nz - 2, 2:g.nr - 2] = True
    for name, bit in (("Ez", abi.FL_EZ_ZERO), ("Er", abi.FL_ER_ZERO), ("Bphi", abi.FL_BPHI_ZERO),
                      ("Ephi", abi.FL_EPHI_ZERO), ("Br", abi.FL_BR_ZERO), ("Bz", abi.FL_BZ_ZERO)):
        sel = interior & ((f & bit) != 0)
        assert sel.any()
        assert np.array_equal(a[name][sel], init[name][sel]), name
    # left/right Neumann rows of Er are zero; axis mirrors
    assert np.array_equal(a["Er"][1:g.nz - 1, 0], -a["Er"][1:g.nz - 1, 1])
    assert np.array_equal(a["Bphi"][1:g.nz, 0], -a["Bphi"][1:g.nz, 1])
    assert np.all(a["Ez"][:, g.nr - 1] == 0)
    assert np.all(a["Ephi"][:, 0] == 0)


# ---------------------------------------------------------------- push and gather
def _one_particle(g, z, r, pz=0.0, pr=0.0, rp=0.0, q=-1e-15):
    gam = np.sqrt(pz * pz + pr * pr + (rp / r if r else 0) ** 2 + 1)
    return {"Q": [q], "z": [z], "r": [r], "pz": [pz], "pr": [pr], "r_pphi": [rp],
            "vz": [K.c_mks * pz / gam], "vr": [K.c_mks * pr / gam], "vphi": [K.c_mks * rp / r / gam if r else 0.0],
            "r0": [r], "z0": [z], "t0": [0.0], "rmax": [r], "rmin": [r], "zmax": [z], "zmin": [z],
            "particleID": [7], "generation": [0], "status": [abi.PART_ACTIVE]}


def test_constant_ez_kick_exact():
    """Uniform constant Ez: each step adds dt*QoMC*Ez to pz exactly (src/advance_particles.c:286)."""
    E0 = 3.0e6
    g = pillbox_grid(L=0.5, nz=101)
    g.constant_fields = (E0, 0, 0, 0, 0, 0)
    o = OracleSim(g, pillbox_flags(g), max_particles=1)
    o.append(_one_particle(g, 0.1, 0.01))
    inc = ((1 * g.dt) * K.QoMC_electron * (E0 + 0.0 * 0 - 0.0 * 0)) / 1.0
    pz = 0.0
    for _ in range(200):
        o.advance_particles()
        pz = pz + inc
        assert o.p["pz"][0] == pz
    assert pz == pytest.approx(200 * g.dt * K.QoMC_electron * E0, rel=1e-12)


def test_uniform_bz_centred_orbit_keeps_radius():
    """Uniform Bz with r*pphi = -QoMC*Bz*r^2: a circular orbit about the axis; r and |p| stay fixed."""
    B0 = 0.05
    g = pillbox_grid()
    g.constant_fields = (0, 0, 0, B0, 0, 0)
    r0 = 0.01
    pphi = -K.QoMC_electron * B0 * r0
    o = OracleSim(g, pillbox_flags(g), max_particles=1)
    o.append(_one_particle(g, 0.0125, r0, rp=r0 * pphi))
    for _ in range(2000):
        o.advance_particles()
    p = o.particles()
    assert abs(p["r"][0] / r0 - 1) < 1e-6
    assert abs(p["r_pphi"][0] / (r0 * pphi) - 1) < 1e-6
    assert abs(p["pr"][0]) < 1e-6 * pphi


@pytest.mark.parametrize("comp", ["Ez", "Er", "Ephi", "Br2", "Bz2", "Bphi2"])
def test_gather_reproduces_linear_fields(comp):
    """Bilinear gather of a field linear in z and r is exact up to rounding, for each stagger
    (src/spiffe.h:60-71): this pins every component's index and fraction rule."""
    g = pillbox_grid(modes=TMTE)
    o = OracleSim(g, pillbox_flags(g))
    half_z = comp in ("Ez", "Br2", "Bphi2")
    half_r = comp in ("Er", "Bz2", "Bphi2")
    zc = g.zmin + g.dz * (np.arange(g.nz + 1) - (0.5 if half_z else 0.0))
    rc = g.dr * (np.arange(g.nr + 1) - (0.5 if half_r else 0.0))
    a, b, c0 = 2.0e6, -3.0e6, 1.0e5
    o.set_field(comp, a * zc[:, None] + b * rc[None, :] + c0)
    rng = np.random.default_rng(2)
    names = {"Ez": 0, "Er": 1, "Ephi": 2, "Bz2": 3, "Br2": 4, "Bphi2": 5}
    for _ in range(200):
        z = g.zmin + g.dz * (3 + rng.random() * (g.nz - 8))
        r = g.dr * (3 + rng.random() * (g.nr - 8))
        got = o.interpolate(z, r)[names[comp]]
        assert got == pytest.approx(a * z + b * r + c0, rel=1e-11, abs=1e-3)


def test_screen_crossing_time_and_radius_straight_line():
    """Field-free straight line: the row's t and r equal the analytic crossing."""
    g = pillbox_grid(L=0.2, nz=401)
    pz, pr = 2.0, 0.05
    z_init, r_init = 0.01, 0.004
    o = OracleSim(g, pillbox_flags(g), max_particles=1)
    o.append(_one_particle(g, z_init, r_init, pz=pz, pr=pr))
    gam = np.sqrt(pz * pz + pr * pr + 1)
    vz, vr = K.c_mks * pz / gam, K.c_mks * pr / gam
    zs = 0.0777
    rows = []
    for step in range(400):
        o.f.time_step = step
        o.f.time = step * g.dt
        rw, _ = o.advance_particles(screens=[(zs, 1)], compat=False)
        rows.extend(rw)
    assert len(rows) == 1
    t_cross = -g.dt + (zs - z_init) / vz       # position z_init is at time 0 - dt
    assert rows[0]["t"] == pytest.approx(t_cross, rel=1e-11)
    assert rows[0]["r"] == pytest.approx(r_init + vr * (t_cross + g.dt), rel=1e-9)
    assert rows[0]["pz"] == pytest.approx(pz, rel=1e-12)


def _reference_compaction(lost, compat):
    """Pure-Python transcription of the loop at src/advance_particles.c:491-518 on a list of ids."""
    ids, is_lost, np_ = list(range(len(lost))), list(lost), len(lost)
    ip = 0
    while ip < np_:
        if is_lost[ip]:
            np_ -= 1
            ids[ip] = ids[np_]
            if not compat:
                is_lost[ip] = is_lost[np_]
            ip -= 1
        ip += 1
    return ids[:np_]


def test_compaction_semantics_brute_force():
    """compat reproduces the reference (np truncated at the first lost index);
    fixed is a swap-remove in which the k-th hole receives the k-th survivor from the end."""
    assert _reference_compaction([0, 1, 0, 0, 1, 0], True) == [0]
    assert _reference_compaction([0, 1, 0, 0, 1, 0], False) == [0, 5, 2, 3]
    assert _reference_compaction([1, 1, 0, 1, 0], False) == [4, 2]
    g = pillbox_grid(L=0.2, nz=401)
    rng = np.random.default_rng(4)
    for compat in (True, False):
        for trial in range(20):
            n = int(rng.integers(1, 40))
            lost = rng.random(n) < 0.3
            o = OracleSim(g, pillbox_flags(g), max_particles=n)
            p = random_particles(g, n, trial)
            p["vz"][:] = 0
            p["vr"][:] = 0
            p["r_pphi"][:] = 0
            p["vphi"][:] = 0
            p["z"][lost] = g.zmax + 0.01       # already outside: lost this step (loss code 1)
            p["particleID"] = np.arange(n, dtype=np.int32)
            o.append(p)
            o.advance_particles(compat=compat)
            assert list(o.particles()["particleID"]) == _reference_compaction(list(lost), compat)


def test_lost_row_on_the_wall():
    """A particle leaving through z = zmax is lost with zLost just inside the last cell."""
    g = pillbox_grid(L=0.05, nz=101)
    o = OracleSim(g, pillbox_flags(g), max_particles=1)
    o.append(_one_particle(g, g.zmax - 0.3 * g.dz, 0.01, pz=5.0))
    _, lost = o.advance_particles(compat=False)
    assert len(lost) == 1
    assert lost[0]["directionCode"] == 2       # GOING_RIGHT
    assert lost[0]["zLost"] == pytest.approx(g.zmax, abs=1.01 * g.dz)
    assert o.b.np == 0


# ---------------------------------------------------------------- deposition
def test_single_particle_deposit_closed_form():
    """One particle: J = Q v / (pi dz dr^2 ((ir+1/2)^2 - (ir-1/2)^2)), 4 Q v/(pi dz dr^2) on axis."""
    g = pillbox_grid(modes=TMTE | abi.SPACE_CHARGE)
    for (z, r) in [(0.0101, 0.0123), (0.0052, 0.0001), (0.02, 0.0301)]:
        o = OracleSim(g, pillbox_flags(g), max_particles=1)
        p = _one_particle(g, z, r, pz=0.3, pr=0.1, rp=r * 0.05)
        o.append(p)
        o.add_particle_currents()
        Q, vz, vr, vphi = p["Q"][0], p["vz"][0], p["vr"][0], p["vphi"][0]
        base = Q / (np.pi * g.dz * g.dr ** 2)
        izh = int((z - g.zmin) / g.dz + 0.5)
        iza = int((z - g.zmin) / g.dz + 1.0)
        irh = int(r / g.dr + 0.5)
        ira = int(r / g.dr + 1.0)
        vol_c = 0.25 if irh == 0 else 2 * irh
        assert o.arr["Jz"][iza, irh] == pytest.approx(base * vz / vol_c, rel=1e-14)
        assert o.arr["Jphi"][izh, irh] == pytest.approx(base * vphi / vol_c, rel=1e-14)
        assert o.arr["Jr"][izh, ira] == pytest.approx(base * vr / (2 * ira - 1), rel=1e-14)
        assert np.count_nonzero(o.arr["Jz"]) == 1 and np.count_nonzero(o.arr["Jr"]) == 1


def test_deposit_charge_flux_invariant():
    """sum J * cell volume = sum Q v over particles away from the top rows."""
    g = pillbox_grid(modes=TMTE | abi.SPACE_CHARGE)
    n = 5000
    p = random_particles(g, n, 9)
    p["r"] = np.minimum(p["r"], g.rmax * 0.9)
    o = OracleSim(g, pillbox_flags(g), max_particles=n)
    o.append(p)
    o.add_particle_currents()
    ir = np.arange(g.nr + 1)
    vc = np.pi * g.dz * g.dr ** 2 * np.where(ir == 0, 0.25, 2.0 * ir)
    vo = np.pi * g.dz * g.dr ** 2 * (2.0 * ir - 1)
    assert np.sum(o.arr["Jz"] * vc) == pytest.approx(np.sum(p["Q"] * p["vz"]), rel=1e-11)
    assert np.sum(o.arr["Jphi"] * vc) == pytest.approx(np.sum(p["Q"] * p["vphi"]), rel=1e-11)
    sel = (p["r"] / g.dr + 1.0).astype(int) >= 1
    assert np.sum(o.arr["Jr"][:, 1:] * vo[1:]) == pytest.approx(np.sum((p["Q"] * p["vr"])[sel]), rel=1e-11)


# ---------------------------------------------------------------- Poisson
def _laplacian(P, dz, dr):
    """Cylindrical 5-point Laplacian in flux form: (1/r) d/dr(r dP/dr) + d2P/dz2, axis via 2 d2P/dr2."""
    L = np.zeros_like(P)
    nzp, nrp = P.shape
    for ir in range(nrp - 1):
        zz = (np.roll(P[:, ir], -1) + np.roll(P[:, ir], 1) - 2 * P[:, ir]) / dz ** 2
        if ir == 0:
            rr = 4 * (P[:, 1] - P[:, 0]) / dr ** 2
        else:
            rp, rm = ir + 0.5, ir - 0.5
            rr = (rp * (P[:, ir + 1] - P[:, ir]) - rm * (P[:, ir] - P[:, ir - 1])) / (ir * dr ** 2)
        L[:, ir] = zz + rr
    return L


def test_sor_recovers_a_known_discrete_solution():
    """Build rhs = L(Psi_true) for a smooth Psi_true held fixed on metal/edge rows; SOR must return Psi_true."""
    g = Grid(nz=41, nr=31, zmin=0.0, zmax=0.04, rmax=0.03, dt=1e-13)
    f = derive_flags(pillbox_marks(g.nz, g.nr))
    rng = np.random.default_rng(5)
    zz = g.zmin + g.dz * np.arange(g.nz + 1)
    rr = g.dr * np.arange(g.nr + 1)
    Pt = np.outer(np.sin(np.pi * zz / 0.04), np.cos(np.pi * rr / 0.06)) + 0.01 * rng.standard_normal(g.shape)
    # held values: metal points (walls iz=0, nz-1; top ir=nr-1) keep Pt; Neumann ends are not updated either
    L = _laplacian(Pt, g.dz, g.dr)
    rhs = np.zeros(g.shape)
    rhs[1:g.nz - 1, :g.nr - 1] = L[1:g.nz - 1, :g.nr - 1]
    o = OracleSim(g, f)
    P0 = Pt.copy()
    P0[1:g.nz - 1, :g.nr - 1] = 0.0
    o.set_field("Psi", P0)
    o.set_field("Q", rhs)
    it, res = o.solve_poisson_cyl("Psi", "Q", 1e-13, 20000)
    assert it > 0
    assert np.abs(o.arr["Psi"] - Pt).max() < 1e-9 * np.abs(Pt).max()
    assert res < 1e-13 * np.abs(rhs[:g.nz, :g.nr - 1]).sum()


def test_omega_schedule_tends_to_optimal_sor():
    """Chebyshev omega -> 2 / (1 + sqrt(1 - rho^2)), the optimal SOR factor, src/poisson.c:121-124."""
    nz, nr, dz, dr = 101, 51, 1e-4, 1e-4
    w = np.zeros(400)
    from oracle.oracle import load
    import ctypes
    load().oracle_poisson_omega_schedule(nz, nr, dz, dr, 400, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    rho = (np.cos(np.pi / nz) + (dz / dr) ** 2 * np.cos(np.pi / nr)) / (1 + (dz / dr) ** 2)
    assert w[0] == 1.0
    assert w[1] == pytest.approx(1 / (1 - rho * rho / 2), rel=1e-15)
    assert w[-1] == pytest.approx(2 / (1 + np.sqrt(1 - rho * rho)), rel=1e-9)


def test_parallel_plate_imposed_field():
    """Plates at 0 and V0 a distance d apart: Psi linear, EzImposed = -V0/d (DC-gun pattern)."""
    nz, nr = 41, 21
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=0.02, rmax=0.01, dt=1e-13)
    m = np.zeros(g.shape, dtype=np.int16)
    m[0, :nr] = abi.FL_IS_SURFACE
    m[nz - 1, :nr] = abi.FL_IS_SURFACE
    f = derive_flags(m)
    o = OracleSim(g, f)
    V0 = 1e5
    psi = np.zeros(g.shape)
    zz = g.zmin + g.dz * np.arange(nz + 1)
    psi[:nz, nr - 1] = V0 * zz[:nz] / g.zmax
    psi[nz - 1, :nr] = V0
    o.set_field("PsiImposed", psi)
    o.setup_poisson_correction(step_interval=0, accuracy=1e-10, maximum_iterations=4000)
    o.compute_imposed_Efield(True)
    P = o.arr["PsiImposed"]
    assert np.abs(P[:nz, :nr] - V0 * zz[:nz, None] / g.zmax).max() < 1e-6 * V0
    Ez = o.arr["EzImposed"][2:nz - 1, :nr - 1]
    assert np.abs(Ez / (-V0 / g.zmax) - 1).max() < 1e-5


def test_gauss_law_charge_of_uniform_density():
    """Er = rho r / (2 eps0) (uniform charge density): the apparent cell charge is rho * cell volume."""
    g = pillbox_grid()
    f = np.zeros(g.shape, dtype=np.int16)
    o = OracleSim(g, f, max_particles=1)
    rho = 1e-6
    r_er = g.dr * (np.arange(g.nr + 1) - 0.5)              # Er stagger
    o.set_field("Er", np.tile(rho * r_er / (2 * K.epsilon_o), (g.nz + 1, 1)))
    g.left_bc = g.right_bc = abi.DIRICHLET
    o.f.left_bc = o.f.right_bc = abi.DIRICHLET
    o.setup_poisson_correction(step_interval=1, accuracy=1e-6, maximum_iterations=2)
    o.add_poisson_correction(False)
    # the Q array after the call is the recomputed apparent charge: check rows away from the top
    ir = np.arange(g.nr + 1)
    vol = np.pi * g.dz * g.dr ** 2 * np.where(ir == 0, 0.25, 2.0 * ir)
    # the correction moved E; compare the Gauss charge of the initial field instead
    o2 = OracleSim(g, f, max_particles=1)
    o2.f.left_bc = o2.f.right_bc = abi.DIRICHLET
    o2.set_field("Er", np.tile(rho * r_er / (2 * K.epsilon_o), (g.nz + 1, 1)))
    o2.setup_poisson_correction(step_interval=1, accuracy=1e-6, maximum_iterations=2,
                                error_charge_threshold=1e30)
    o2.add_poisson_correction(False)    # threshold not met -> Q holds the apparent charge, no solve
    Q = o2.arr["Q"][1:g.nz - 1, :g.nr - 2]
    assert np.abs(Q / (rho * vol[None, :g.nr - 2]) - 1).max() < 1e-10


# ---------------------------------------------------------------- antenna
def test_antenna_waveform_linear_interpolation():
    from oracle.oracle import load
    import ctypes
    lib = load()
    t = np.array([0.0, 1e-9, 2e-9, 5e-9])
    W = np.array([0.0, 0.5, 1.0, 1.0])
    li = ctypes.c_int64(0)
    err = ctypes.c_int32(0)
    dp = ctypes.POINTER(ctypes.c_double)
    f, I = 1e9, 2.0
    for time in [0.25e-9, 0.5e-9, 1.5e-9, 3e-9, 6e-9]:
        J = lib.oracle_antenna_current(t.ctypes.data_as(dp), W.ctypes.data_as(dp), 4, ctypes.byref(li), I, f, 0.3,
                                       time, ctypes.byref(err))
        env = np.interp(time, t, W)
        assert err.value == 0
        assert J == pytest.approx(env * I * np.sin(2 * np.pi * f * time + 0.3), rel=1e-12, abs=1e-15)


# ---------------------------------------------------------------- hand-drawn lattices: loss codes, gather
# index fix-ups, loss coordinates, charge removal, field correction


def _lattice_sim(nz, nr, marks, modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP, d=1e-3, cap=0):
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=d * (nz - 1), rmax=d * (nr - 1), dt=courant_dt(d, d, 0.9), modes=modes)
    return g, OracleSim(g, c_derive_flags(marks), max_particles=cap)


def test_loss_codes_hand_lattice():
    """checkLossCode (src/advance_particles.c:31-65) on a hand-drawn 6 x 5 lattice:
    3 inside metal; for a surface point 4 when iz+1 > nz, else 5 when ir+1 > nr,
    else 6 when the three points up and to the right are all metal; 0 otherwise."""
    import ctypes
    S, I = abi.FL_IS_SURFACE, abi.FL_IS_INSIDE
    nz, nr = 5, 4
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[1, 1] = I
    m[2, 1], m[3, 2], m[2, 2], m[3, 1] = S, S, I, S
    m[5, 0], m[5, 4], m[0, 4] = S, S, S
    m[4, 0], m[5, 1], m[4, 1] = S, I, S
    g, o = _lattice_sim(nz, nr, m)
    expected = {(1, 1): 3, (2, 1): 6, (2, 2): 3, (3, 2): 0, (3, 1): 0, (5, 0): 4, (5, 4): 4, (0, 4): 5,
                (4, 0): 6, (5, 1): 3, (4, 1): 0}
    for iz in range(nz + 1):
        for ir in range(nr + 1):
            code = o.lib.oracle_check_loss_code(iz, ir, ctypes.byref(o.f))
            assert code == expected.get((iz, ir), 0), (iz, ir, code)


def _labels(g, base):
    iz, ir = np.meshgrid(np.arange(g.nz + 1), np.arange(g.nr + 1), indexing="ij")
    return base + 100.0 * iz + ir + 1.0


def _frac(x):
    i = int(x)
    return i, x - i


@pytest.mark.parametrize("vacuum_flag", [0, abi.FL_IS_VACUUM])
def test_gather_ez_next_to_a_constant_z_wall(vacuum_flag):
    """The Ez gather's index shifts beside a wall along z = const
    (src/advance_particles.c:683-699): when both rows of the cell at iz have Er
    zeroed, iz moves to iz1 if the column before holds any flag (vacuum drawn as
    FL_IS_VACUUM counts), else iz1 falls back to iz; when only iz1's column has
    Er zeroed, iz1 falls back to iz.  Ez holds a distinct label per point, so the
    gathered value names the points used."""
    S, I = abi.FL_IS_SURFACE, abi.FL_IS_INSIDE
    nz, nr = 10, 8
    m = np.full((nz + 1, nr + 1), vacuum_flag, dtype=np.int16)
    m[6, :] = S
    m[7:, :] = I
    g, o = _lattice_sim(nz, nr, m)
    E = _labels(g, 0.0)
    o.set_field("Ez", E)
    d = g.dz

    def bilinear(iz, iz1, ir, ir1, fl, fr, fa, fb):
        return fb * (fl * E[iz, ir] + fr * E[iz1, ir]) + fa * (fl * E[iz, ir1] + fr * E[iz1, ir1])

    ir, fa = _frac(2.3e-3 / d)
    # cell column 6 has Er zeroed in both rows: iz = 6
    iz, fr = _frac(5.6e-3 / d + 0.5)
    assert iz == 6
    got = o.interpolate(5.6e-3, 2.3e-3)[0]
    if vacuum_flag:   # column 5 carries a flag: iz moves to iz1 = 7
        want = bilinear(7, 7, ir, ir + 1, 1 - fr, fr, fa, 1 - fa)
    else:             # column 5 is bare: iz1 falls back to 6
        want = bilinear(6, 6, ir, ir + 1, 1 - fr, fr, fa, 1 - fa)
    assert got == pytest.approx(want, rel=1e-12)
    # iz = 5 with column 6 zeroed: iz1 falls back to 5
    iz, fr = _frac(4.8e-3 / d + 0.5)
    assert iz == 5
    assert o.interpolate(4.8e-3, 2.3e-3)[0] == pytest.approx(bilinear(5, 5, ir, ir + 1, 1 - fr, fr, fa, 1 - fa),
                                                              rel=1e-12)
    # far from the wall: plain bilinear over columns 2 and 3
    iz, fr = _frac(2.2e-3 / d + 0.5)
    assert o.interpolate(2.2e-3, 2.3e-3)[0] == pytest.approx(bilinear(2, 3, ir, ir + 1, 1 - fr, fr, fa, 1 - fa),
                                                              rel=1e-12)


def test_gather_er_next_to_a_constant_r_wall():
    """The Er gather's index shifts beside a wall along r = const
    (src/advance_particles.c:744-747): with Ez zeroed on both columns of row ir
    the gather moves ir to ir+1; with Ez zeroed on both columns of row ir+1 it
    uses row ir twice; a column at iz = 0 never has Ez zeroed."""
    S, I = abi.FL_IS_SURFACE, abi.FL_IS_INSIDE
    nz, nr = 10, 8
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[:, 5] = S
    m[:, 6:] = I
    g, o = _lattice_sim(nz, nr, m)
    E = _labels(g, 10000.0)
    o.set_field("Er", E)
    d = g.dr

    def bilinear(iz, ir, ir1, fl, fr, fa, fb):
        return fl * (fb * E[iz, ir] + fa * E[iz, ir1]) + fr * (fb * E[iz + 1, ir] + fa * E[iz + 1, ir1])

    iz, fr = _frac(3.4e-3 / d)
    ir, fa = _frac(4.3e-3 / d + 0.5)        # row 4, row 5 above has Ez zeroed: row 4 twice
    assert (iz, ir) == (3, 4)
    assert o.interpolate(3.4e-3, 4.3e-3)[1] == pytest.approx(bilinear(3, 4, 4, 1 - fr, fr, fa, 1 - fa), rel=1e-12)
    ir, fa = _frac(4.6e-3 / d + 0.5)        # row 5 has Ez zeroed: rows 6 and 6
    assert ir == 5
    assert o.interpolate(3.4e-3, 4.6e-3)[1] == pytest.approx(bilinear(3, 6, 6, 1 - fr, fr, fa, 1 - fa), rel=1e-12)
    iz, fr = _frac(0.4e-3 / d)              # iz = 0: point (0, 5) has no Ez zero, no shift
    assert o.interpolate(0.4e-3, 4.6e-3)[1] == pytest.approx(bilinear(0, 5, 6, 1 - fr, fr, fa, 1 - fa), rel=1e-12)


def _loss_coords(o, z0, r0, z1, r1):
    import ctypes
    zr, rr, code = ctypes.c_double(), ctypes.c_double(), ctypes.c_short()
    o.lib.oracle_find_loss_coordinates(ctypes.byref(zr), ctypes.byref(rr), ctypes.byref(code), z0, r0, z1, r1,
                                       ctypes.byref(o.f))
    return zr.value, rr.value, code.value


def test_loss_coordinates_by_direction():
    """findLossCoordinates (src/secondaryEmission.c:239-372): the direction code
    (1 left, 2 right, 3 up, 4 down) from the cells left and entered, and the
    point where the straight step meets the cell face it names (a 1e-12 cell
    inside it).  Expected points are intersections worked out by hand."""
    S, I = abi.FL_IS_SURFACE, abi.FL_IS_INSIDE
    nz, nr = 10, 8
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[4, 3] = S
    g, o = _lattice_sim(nz, nr, m)
    mm = 1e-3
    # no motion: code 1 at the start
    assert _loss_coords(o, 3.3 * mm, 2.2 * mm, 3.3 * mm, 2.2 * mm) == (3.3 * mm, 2.2 * mm, 1)
    # same row, moving right: face z = 4 mm, r on the line there = 2.5 + 0.2/0.7*0.5
    z, r, c = _loss_coords(o, 3.5 * mm, 2.5 * mm, 4.2 * mm, 2.7 * mm)
    assert c == 2 and z == pytest.approx(4 * mm, abs=1e-14) and r == pytest.approx((2.5 + 0.5 * 0.2 / 0.7) * mm, rel=1e-10)
    # same column, moving down: face r = 2 mm, z = 3.3 + 0.3 * 0.2/0.5
    z, r, c = _loss_coords(o, 3.3 * mm, 2.2 * mm, 3.6 * mm, 1.7 * mm)
    assert c == 4 and r == pytest.approx(2 * mm, abs=1e-14) and z == pytest.approx(3.42 * mm, rel=1e-10)
    # same column, moving up: face r = 3 mm, z = 3.3 + 0.3 * 0.8/0.9
    z, r, c = _loss_coords(o, 3.3 * mm, 2.2 * mm, 3.6 * mm, 3.1 * mm)
    assert c == 3 and r == pytest.approx(3 * mm, abs=1e-14) and z == pytest.approx((3.3 + 0.3 * 0.8 / 0.9) * mm, rel=1e-10)
    # diagonal into the metal point (4, 3): right, face z = 4 mm, r = 3.0 mm
    z, r, c = _loss_coords(o, 3.8 * mm, 2.8 * mm, 4.3 * mm, 3.3 * mm)
    assert c == 2 and z == pytest.approx(4 * mm, abs=1e-14) and r == pytest.approx(3.0 * mm, rel=1e-10)
    # diagonal into vacuum: no branch names a face, the code stays 0 and the
    # up-face formula applies: r = 3 mm, z on the line = 5.0 mm
    z, r, c = _loss_coords(o, 4.8 * mm, 2.8 * mm, 5.3 * mm, 3.3 * mm)
    assert c == 0 and r == pytest.approx(3 * mm, abs=1e-14) and z == pytest.approx(5.0 * mm, rel=1e-10)
    # leaving through z < zmin: left, face z = 0, r = 2.5 + 0.7 * 0.5/0.8
    z, r, c = _loss_coords(o, 0.5 * mm, 2.5 * mm, -0.3 * mm, 3.2 * mm)
    assert c == 1 and z == pytest.approx(0.0, abs=1e-14) and r == pytest.approx((2.5 + 0.7 * 0.5 / 0.8) * mm, rel=1e-10)
    # crossing the axis within a column: down, r returned as |r| at the axis face
    z, r, c = _loss_coords(o, 3.3 * mm, 0.4 * mm, 3.6 * mm, -0.3 * mm)
    assert c == 4 and r == pytest.approx(0.0, abs=1e-14) and z == pytest.approx((3.3 + 0.3 * 0.4 / 0.7) * mm, rel=1e-10)


def _particles_at(zs, rs, q):
    n = len(zs)
    z, r = np.asarray(zs, float), np.asarray(rs, float)
    return {"Q": np.full(n, q), "z": z, "r": r, "pz": np.zeros(n), "pr": np.zeros(n), "r_pphi": np.zeros(n),
            "vz": np.zeros(n), "vr": np.zeros(n), "vphi": np.zeros(n), "r0": r.copy(), "z0": z.copy(),
            "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(), "zmax": z.copy(), "zmin": z.copy(),
            "particleID": np.arange(n, dtype=np.int32), "generation": np.zeros(n, dtype=np.int16),
            "status": np.full(n, abi.PART_ACTIVE, dtype=np.uint32)}


def test_charge_removal_lands_on_the_nearest_point():
    """remove_particle_charge (src/poisson_correction.c:184-243) without
    smearing: with zero fields the Gauss-law charge is zero, and each particle's
    charge is taken from the nearest grid point (iz = (z-zmin)/dz + 0.5, ir =
    r/dr + 0.5), except at metal points and at the Neumann rows (iz = 0 on the
    left, iz >= nz-1 on the right, ir >= nr-1 on top); Qreal is the charge
    removed.  A threshold above the error charge stops the correction there."""
    S = abi.FL_IS_SURFACE
    nz, nr = 10, 8
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[5, 2] = S
    g, o = _lattice_sim(nz, nr, m, modes=abi.TM_FIELDS | abi.SPACE_CHARGE, cap=16)
    mm = 1e-3
    q = -2e-15
    cases = [((3.4, 2.2), (3, 2)), ((3.6, 2.6), (4, 3)), ((7.49, 0.1), (7, 0)), ((4.9, 2.1), None),  # metal (5,2)
             ((0.3, 3.0), None), ((8.6, 3.0), None), ((4.0, 6.6), None)]                            # Neumann rows
    o.append(_particles_at([c[0][0] * mm for c in cases], [c[0][1] * mm for c in cases], q))
    o.setup_poisson_correction(step_interval=1, accuracy=1e-6, maximum_iterations=10, error_charge_threshold=1.0)
    assert o.add_poisson_correction(True) == -2
    Q = o.field("Q")
    want = np.zeros(g.shape)
    for _, cell in cases:
        if cell:
            want[cell] -= q
    assert np.array_equal(Q, want)
    assert o.pc.Qreal == 3 * q
    assert o.pc.Qerror == -3 * q


def test_field_correction_restores_gauss_law():
    """The correction subtracts the gradient of the error potential from E
    (src/poisson_correction.c:289-326), so that the charge Gauss's law finds in
    E afterwards (recomputed, :332-439) matches the particles' charge: the
    apparent minus particle charge left in Q shrinks from the initial error to
    the solver's residual.  A sign slip in the gradient would double the error
    and an index slip would leave it of the same size."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    o = OracleSim(g, f, max_particles=4000)
    rng = np.random.default_rng(5)
    for name in ("Ez", "Er"):
        o.set_field(name, rng.standard_normal(g.shape) * 1e3)
    o.append(random_particles(g, 4000, 6, flags=f, q=-1e-14))
    probe = OracleSim(g, f, max_particles=4000)
    for name in ("Ez", "Er"):
        probe.set_field(name, o.field(name))
    probe.append(o.particles())
    probe.setup_poisson_correction(step_interval=1, error_charge_threshold=1e30)
    probe.add_poisson_correction(True)
    before = probe.field("Q")
    o.setup_poisson_correction(step_interval=1, accuracy=1e-11, maximum_iterations=20000)
    assert o.add_poisson_correction(True) > 0
    after = o.field("Q")
    interior = (f[:g.nz, :g.nr - 1] & abi.FL_IS_METAL) == 0
    interior[0, :] = interior[g.nz - 1, :] = False
    b, a = np.abs(before[:g.nz, :g.nr - 1][interior]), np.abs(after[:g.nz, :g.nr - 1][interior])
    assert b.max() > 0
    assert a.max() < 1e-6 * b.max()


# ---------------------------------------------------------------- area-weighted (smeared) deposition
SMEARS = {"radial": abi.RADIAL_SMEAR, "longitudinal": abi.LONGIT_SMEAR,
          "both": abi.RADIAL_SMEAR | abi.LONGIT_SMEAR}


def _interior_particles(g, n, seed):
    """Particles away from the z ends and the top rows, r >= dr/2 (the Jr row-0 rule
    puts every current of a particle below dr/2 on row 1)."""
    rng = np.random.default_rng(seed)
    p = random_particles(g, n, seed)
    p["z"] = g.zmin + g.dz * (2.0 + rng.random(n) * (g.nz - 5))
    p["r"] = g.dr * (0.5 + rng.random(n) * (g.nr - 4))
    return p


@pytest.mark.parametrize("smear", sorted(SMEARS))
def test_smeared_deposit_conserves_current_and_centres_it_on_the_particle(smear):
    """Area weighting (src/space_charge.c, FL_RADIAL_SMEAR / FL_LONGIT_SMEAR) splits
    each particle's current between the two nearest nodes in each smeared
    direction with weights that sum to one and whose centroid is the particle:
    sum J*vol = sum Q v, and sum J*vol*node = sum Q v * position along each
    smeared direction, with Jz nodes at z = (iz - 1/2) dz and Jr/Jphi nodes at
    iz dz, Jz/Jphi rows at r = ir dr and Jr rows at (ir - 1/2) dr (vol: the
    annulus areas of the staggered rows).  A swapped weight, a shifted node or a
    wrong volume factor moves a centroid or loses current."""
    g = pillbox_grid(modes=TMTE | abi.SPACE_CHARGE | SMEARS[smear])
    n = 3000
    p = _interior_particles(g, n, 21)
    o = OracleSim(g, pillbox_flags(g), max_particles=n)
    o.append(p)
    o.add_particle_currents()
    ir = np.arange(g.nr + 1, dtype=float)
    iz = np.arange(g.nz + 1, dtype=float)[:, None]
    base = np.pi * g.dz * g.dr ** 2
    vc = base * np.where(ir == 0, 0.25, 2.0 * ir)       # Jz, Jphi rows
    vo = base * (2.0 * ir - 1)                           # Jr rows (ir >= 1)
    x = (p["z"] - g.zmin) / g.dz
    y = p["r"] / g.dr
    for name, v, vol, zn, rn in (("Jz", p["vz"], vc, iz - 0.5, ir), ("Jphi", p["vphi"], vc, iz, ir),
                                 ("Jr", p["vr"], vo, iz, ir - 0.5)):
        Jv = o.arr[name] * vol
        Qv = p["Q"] * v
        assert np.sum(Jv) == pytest.approx(np.sum(Qv), rel=1e-11), name
        if SMEARS[smear] & abi.LONGIT_SMEAR:
            assert np.sum(Jv * zn) == pytest.approx(np.sum(Qv * x), rel=1e-11), name + " z centroid"
        if SMEARS[smear] & abi.RADIAL_SMEAR:
            assert np.sum(Jv * rn) == pytest.approx(np.sum(Qv * y), rel=1e-11), name + " r centroid"


@pytest.mark.parametrize("smear", sorted(SMEARS))
def test_smeared_charge_removal_centres_the_charge_on_the_particle(smear):
    """remove_particle_charge with smearing (src/poisson_correction.c:188-239):
    with zero fields Q holds minus the removed charge; away from metal and the
    Neumann rows each particle's charge comes off in full, centred on the
    particle in each smeared direction (nodes at iz dz, ir dr), and Qreal is the
    charge removed."""
    nz, nr = 24, 16
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    g, o = _lattice_sim(nz, nr, m, modes=abi.TM_FIELDS | abi.SPACE_CHARGE | SMEARS[smear], cap=400)
    rng = np.random.default_rng(4)
    n = 400
    zs = g.zmin + g.dz * (1.5 + rng.random(n) * (nz - 4))
    rs = g.dr * rng.random(n) * (nr - 3)
    q = -1e-15 * (1 + rng.random(n))
    p = _particles_at(zs, rs, 0.0)
    p["Q"] = q
    o.append(p)
    o.setup_poisson_correction(step_interval=1, accuracy=1e-6, maximum_iterations=10, error_charge_threshold=1.0)
    assert o.add_poisson_correction(True) == -2
    removed = -o.field("Q")
    assert np.sum(removed) == pytest.approx(np.sum(q), rel=1e-12)
    assert o.pc.Qreal == pytest.approx(np.sum(q), rel=1e-12)
    x, y = (zs - g.zmin) / g.dz, rs / g.dr
    izg = np.arange(nz + 1, dtype=float)[:, None]
    irg = np.arange(nr + 1, dtype=float)[None, :]
    if SMEARS[smear] & abi.LONGIT_SMEAR:
        assert np.sum(removed * izg) == pytest.approx(np.sum(q * x), rel=1e-12)
    if SMEARS[smear] & abi.RADIAL_SMEAR:
        assert np.sum(removed * irg) == pytest.approx(np.sum(q * y), rel=1e-12)


# ---------------------------------------------------------------- Poisson guesses
def _guess_case(code):
    """A lattice with metal points (some on the axis), random fields and a few
    particles; one correction with the given guess and a short solve.  The SOR
    never updates metal points or the row ir = nr-1, so Psi there keeps the
    guess of src/poisson_correction.c:564-615."""
    nz, nr = 30, 14
    S = abi.FL_IS_SURFACE
    m = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    m[10:13, 0:3] = S
    m[20, 5] = S
    g, o = _lattice_sim(nz, nr, m, modes=abi.TM_FIELDS | abi.SPACE_CHARGE, cap=50)
    rng = np.random.default_rng(8)
    fields = {k: rng.standard_normal(g.shape) * 1e3 for k in ("Ez", "Er")}
    for k, v in fields.items():
        o.set_field(k, v)
    p = _particles_at(g.zmin + g.dz * (3 + rng.random(50) * 20), g.dr * rng.random(50) * 8, -3e-16)
    o.append(p)
    # the Gauss-law charge before any particle charge comes off: the same
    # fields with a correction that stops at its error-charge test
    probe = _lattice_sim(nz, nr, m, modes=abi.TM_FIELDS | abi.SPACE_CHARGE, cap=50)[1]
    for k, v in fields.items():
        probe.set_field(k, v)
    probe.setup_poisson_correction(step_interval=1, error_charge_threshold=1e30)
    probe.add_poisson_correction(False)
    gauss = probe.field("Q")
    o.setup_poisson_correction(step_interval=1, accuracy=1e-3, maximum_iterations=3, guess_code=code)
    o.add_poisson_correction(True)
    keep = np.zeros(g.shape, dtype=bool)
    keep[:nz, nr - 1] = True
    keep[:nz, :nr] |= (m[:nz, :nr] & abi.FL_IS_METAL) != 0
    return g, o, gauss, keep


def test_line_charge_guess():
    """guess_type "line-charge": Psi = Qerror/(2 pi eps0) ln r on every row,
    r = ir dr, and dr/2 on the axis row."""
    g, o, _, keep = _guess_case(1)
    psi = o.field("Psi")
    assert o.pc.Qerror != 0
    r = g.dr * np.where(np.arange(g.nr + 1) == 0, 0.5, np.arange(g.nr + 1))
    want = np.broadcast_to(o.pc.Qerror / (2 * math.pi * K.epsilon_o) * np.log(r)[None, :], psi.shape)
    assert keep[11, 0] and keep[5, g.nr - 1]
    np.testing.assert_allclose(psi[keep], want[keep], rtol=1e-14)


def test_point_charge_guess():
    """guess_type "point-charge": Psi = Qerror/(4 pi eps0 d), d the distance from
    (iz dz + zmin, (ir + 1/2) dr) to (c, 0), where c is the reference's
    "centroid": the Gauss-law charges times their z, summed without dividing by
    the total (src/poisson_correction.c:176)."""
    g, o, gauss, keep = _guess_case(2)
    psi = o.field("Psi")
    c = math.fsum((gauss[iz, ir] * (iz * g.dz + g.zmin) for ir in range(g.nr) for iz in range(g.nz)))
    assert c != 0 and o.pc.Qerror != 0
    izg = np.arange(g.nz + 1)[:, None] * g.dz + g.zmin
    irg = (np.arange(g.nr + 1)[None, :] + 0.5) * g.dr
    want = o.pc.Qerror / (4 * math.pi * K.epsilon_o * np.sqrt((izg - c) ** 2 + irg ** 2))
    np.testing.assert_allclose(psi[keep], want[keep], rtol=1e-12)


def test_j_filter_is_an_exponential_average():
    """integrate's J_filter_multiplier m: J_n = m J_(n-1) + (1-m) L for a line
    current L added every step, so J_n = L (1 - m^n) from J_0 = 0; scaling in
    the wrong place (before the lines, or not at all) gives L (1-m^n)/(1-m) or
    n L instead."""
    g = pillbox_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    o = OracleSim(g, pillbox_flags(g), max_particles=1)
    m, L = 0.35, 4.0e3
    lines = ((0, 5, 10, 20, L),)
    for step in range(1, 9):
        o.step(lines=lines, j_filter=m)
        J = o.arr["Jz"]
        want = L * (1 - m ** step)
        assert J[10:21, 5] == pytest.approx(np.full(11, want), rel=1e-13)
        assert np.count_nonzero(J) == 11
