"""The oracle's solenoid field (src/solenoid.c, src/Bsolenoid.c) pinned by
scipy's elliptic integrals, the closed-form axis field of a ring, Biot-Savart
quadrature off the axis, superposition and div B = 0."""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.special import ellipe, ellipk

from oracle import oracle as O
from tests import _support as S

MU0 = 4e-7 * math.pi


def ke(k):
    K, E = ctypes.c_double(), ctypes.c_double()
    O.load().oracle_ellip_ke(k, ctypes.byref(K), ctypes.byref(E))
    return K.value, E.value


def ring(rho, z, R):
    br, bz = ctypes.c_double(), ctypes.c_double()
    O.load().oracle_b_ring(ctypes.byref(br), ctypes.byref(bz), rho, z, R)
    return br.value, bz.value


def sol(rho, z, R, z0, z1, n, sym):
    br, bz = ctypes.c_double(), ctypes.c_double()
    O.load().oracle_b_solenoid(ctypes.byref(br), ctypes.byref(bz), rho, z, R, z0, z1, n, sym)
    return br.value, bz.value


def test_elliptic_integrals_match_scipy():
    for k in np.concatenate([np.linspace(0, 0.99, 200), [0.999, 0.99999]]):
        K, E = ke(k)
        assert K == pytest.approx(ellipk(k * k), rel=4e-15)
        assert E == pytest.approx(ellipe(k * k), rel=4e-15)
    assert ke(0.0) == (math.pi / 2, math.pi / 2)


def test_ring_axis_field_is_the_textbook_value():
    R = 0.05
    for z in [-0.2, -0.03, 0.0, 0.01, 0.5]:
        br, bz = ring(0.0, z, R)
        assert br == 0
        assert bz == pytest.approx(MU0 * R * R / (2 * (R * R + z * z) ** 1.5), rel=1e-14)


def biot_savart(rho, z, R):
    # field at (rho, 0, z) of 1 A in the ring x^2+y^2 = R^2 at z = 0
    def dB(phi):
        lx, ly = -R * math.sin(phi), R * math.cos(phi)          # dl/dphi
        rx, ry, rz = rho - R * math.cos(phi), -R * math.sin(phi), z
        d3 = (rx * rx + ry * ry + rz * rz) ** 1.5
        return (ly * rz / d3, (lx * ry - ly * rx) / d3)         # (dB_x, dB_z) up to mu0/4pi
    bx = integrate.quad(lambda p: dB(p)[0], 0, 2 * math.pi, epsabs=0, epsrel=1e-11, limit=200)[0]
    bz = integrate.quad(lambda p: dB(p)[1], 0, 2 * math.pi, epsabs=0, epsrel=1e-11, limit=200)[0]
    return 1e-7 * bx, 1e-7 * bz


@pytest.mark.parametrize("rho,z", [(0.01, 0.0), (0.03, 0.02), (0.049, -0.004), (0.08, 0.05), (0.02, -0.3)])
def test_ring_off_axis_matches_biot_savart(rho, z):
    R = 0.05
    br, bz = ring(rho, z, R)
    qr, qz = biot_savart(rho, z, R)
    assert br == pytest.approx(qr, rel=1e-9, abs=1e-18)
    assert bz == pytest.approx(qz, rel=1e-9)


def test_solenoid_is_a_sum_of_rings_with_mirror_images():
    R, z0, z1, n = 0.04, 0.05, 0.09, 5
    for rho, z in [(0.0, 0.07), (0.01, 0.0), (0.03, -0.06)]:
        ebr = ebz = 0.0
        for i in range(n):
            zc = z0 + i * (z1 - z0) / (n - 1)
            a, b = ring(rho, z - zc, R)
            c, d = ring(rho, z + zc, R)
            ebr, ebz = ebr + a + c, ebz + b + d
        br, bz = sol(rho, z, R, z0, z1, n, 1)
        assert br == pytest.approx(ebr, rel=1e-13, abs=1e-20) and bz == pytest.approx(ebz, rel=1e-13)
        br2, bz2 = sol(rho, z, R, z0, z1, n, 0)
        br3, bz3 = sol(rho, z, R, z0, z1, n, -1)
        assert br2 - (br - br2) == pytest.approx(br3, rel=1e-12, abs=1e-18)


def gun():
    g = S.gun_grid(modes=S.TMTE)
    return g, O.OracleSim(g, S.gun_flags(g), max_particles=1)


def test_solenoid_grid_scales_to_bz_peak_and_is_divergence_free():
    g, o = gun()
    cur = o.define_solenoid(radius=0.08, z_start=0.09, z_end=0.13, turns=40, current=100.0, Bz_peak=0.25)
    assert cur == 100.0
    Bz, Br = o.field("BzImposed"), o.field("BrImposed")
    assert Bz[:, 0].max() == pytest.approx(0.25, rel=2e-16)
    assert np.all(Br[:, 0] == 0)
    # div B = (1/r) d(r Br)/dr + dBz/dz ~ 0 at second order in the interior
    z = g.zmin + g.dz * np.arange(g.nz + 1)
    r = g.dr * np.arange(g.nr + 1)
    iz, ir = slice(2, g.nz - 2), slice(2, g.nr - 2)
    rBr = Br * r[None, :]
    div = (rBr[iz, 3:g.nr - 1] - rBr[iz, 1:g.nr - 3]) / (2 * g.dr) / r[None, ir] + \
          (Bz[3:g.nz - 1, ir] - Bz[1:g.nz - 3, ir]) / (2 * g.dz)
    scale = np.abs(Bz).max() / 0.01
    assert np.abs(div).max() < 2e-3 * scale


def test_bucking_coil_cancels_the_axis_field_at_z_buck():
    g, o = gun()
    o.define_solenoid(radius=0.08, z_start=0.09, z_end=0.13, turns=20, current=300.0)
    zb = 0.0
    iz = int((zb - g.zmin) / g.dz)
    before = o.field("BzImposed")[iz, 0]
    cur = o.define_solenoid(radius=0.06, z_start=-0.03, z_end=-0.01, turns=8, bucking=1, z_buck=zb)
    assert cur < 0 and before > 0
    assert abs(o.field("BzImposed")[iz, 0]) < 1e-12 * before


def test_gaussian_extension_axis_profile_and_first_order_br():
    g, o = gun()
    o.define_solenoid(gaussian=1, Bz_peak=0.25, gaussian_center=0.20, gaussian_sigma=0.05)
    Bz, Br = o.field("BzImposed"), o.field("BrImposed")
    z = g.zmin + g.dz * np.arange(g.nz + 1)
    r = g.dr * np.arange(g.nr + 1)
    G = 0.25 * np.exp(-(z - 0.2) ** 2 / (2 * 0.05 ** 2))
    assert np.allclose(Bz[:, 0], G, rtol=1e-14)
    assert np.allclose(Bz[:, -1], G, rtol=1e-14)        # first order: Bz(r) = Bz(0)
    dG = -(z - 0.2) / 0.05 ** 2 * G
    assert np.allclose(Br, -(r[None, :] / 2) * dG[:, None], rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("kw,msg", [(dict(radius=0.0), "radius"), (dict(radius=0.1, z_start=1, z_end=0), "z_start"),
                                    (dict(radius=0.1, turns=0), "turns"), (dict(radius=0.1, symmetry=2), "symmetry")])
def test_solenoid_errors(kw, msg):
    g, o = gun()
    with pytest.raises(ValueError, match=msg):
        o.define_solenoid(**kw)
