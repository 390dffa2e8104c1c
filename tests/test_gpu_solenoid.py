"""Imposed solenoid fields computed on the device (one thread per grid point,
exact ring fields) against the oracle, bit for bit, and a beam pushed through
them."""
import numpy as np
import pytest

from _support import TMTE, gun_flags, gun_grid, random_fields, random_particles
from spiffe_amd import abi
from test_gpu_parity import assert_fields, assert_particles, ctx, load_fields, pair  # noqa: F401

pytestmark = pytest.mark.gpu

SOL = ["BzImposed", "BrImposed"]


def test_coil_fields_scaling_limit_and_bucking(ctx):
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    o, d = pair(ctx, g, f)
    coil = dict(radius=0.08, z_start=0.02, z_end=0.13, turns=37, current=120.0, symmetry=1,
                evaluation_radius_limit=0.031, Bz_peak=0.3)
    assert o.define_solenoid(**coil) == d.define_solenoid(**coil)
    assert_fields(o, d, SOL)
    buck = dict(radius=0.05, z_start=-0.05, z_end=-0.01, turns=9, bucking=1, z_buck=0.004)
    co, cd = o.define_solenoid(**buck), d.define_solenoid(**buck)
    assert np.float64(co).view(np.int64) == np.float64(cd).view(np.int64)
    assert_fields(o, d, SOL)
    d.close()


def test_gaussian_solenoid_and_push(ctx):
    g = gun_grid(modes=TMTE)
    f = gun_flags(g)
    n = 3000
    o, d = pair(ctx, g, f, cap=n)
    sol = dict(gaussian=1, Bz_peak=0.25, gaussian_center=0.20, gaussian_sigma=0.05)
    o.define_solenoid(**sol)
    d.define_solenoid(**sol)
    assert_fields(o, d, SOL)
    load_fields(o, d, random_fields(g, 51, scale=5e5))
    p = random_particles(g, n, 52, flags=f, pz_scale=0.3)
    o.append(p)
    d.append(p)
    for _ in range(60):
        o.step()
        d.step()
    assert_particles(o, d)
    d.close()
