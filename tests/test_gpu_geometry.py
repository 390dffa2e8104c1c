"""A run built from a boundary file: the DC-gun deck drawn by the library, its
imposed field solved on the device, and a flat-top pulse emitted, pushed and
counted -- bit for bit against the oracle."""
import math

import numpy as np
import pytest

from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd import workloads as W
from test_gpu_parity import assert_fields, assert_particles, assert_rows, ctx, pair  # noqa: F401

pytestmark = pytest.mark.gpu


def test_dc_gun_deck_field_emission_and_push(ctx):
    g = W.dc_gun_grid(dz=2e-4)                    # 151 x 61 lines
    f, psi = W.dc_gun_geometry(g)
    screens = [(0.008, 1)]
    o, d = pair(ctx, g, f, cap=30000, screens=screens, compat=False)
    o.set_field("PsiImposed", psi)
    d.upload("PsiImposed", psi)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=0, accuracy=1e-10, maximum_iterations=6000)
    o.compute_imposed_Efield(True)
    d.compute_imposed_Efield(True)
    assert_fields(o, d, ["PsiImposed", "EzImposed", "ErImposed"])
    # mid-gap on the axis the graded gap is close to a plane gap: Ez = -V/d
    Ez = d.download("EzImposed")
    iz = int(0.010 / g.dz + 0.5)
    assert Ez[iz, 0] == pytest.approx(-1e5 / 0.020, rel=0.05)
    # 10 ps flat-top (0.5 ps edges) of 50 pC from the 3 mm cathode, 20 k macroparticles
    Q, N, tf = 50e-12, 20000, 10e-12
    area = math.pi * 3e-3 ** 2
    J = Q / tf / area
    prof = (np.array([0.0, 0.5e-12, 10.5e-12, 11e-12]), np.array([0.0, 1.0, 1.0, 0.0]))
    kw = dict(z_position=0.0, outer_radius=3e-3, current_density=J, electrons_per_macroparticle=Q / K.e_mks / N,
              initial_pz=1e-3, zoned_emission=0, random_number_seed=12345)
    assert o.define_cathode(profile=prof, **kw) == d.define_cathode(profile=prof, **kw) == [0]
    rows_o, rows_d, lost_o, lost_d = [], [], [], []
    for step in range(450):
        ro, lo = o.step(screens=screens, compat=False)
        d.step()
        rd, ld = d.drain()
        rows_o.append(ro), rows_d.append(rd), lost_o.append(lo), lost_d.append(ld)
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
    emitted = d.next_particle_id() - 1
    assert 1.0 * N < emitted < 1.1 * N, emitted     # 10.5 ps of full current
    rows = np.concatenate(rows_d)
    assert len(rows) == emitted               # the whole pulse crosses the screen
    assert_particles(o, d)
    assert_rows(np.concatenate(rows_o), rows, np.concatenate(lost_o), np.concatenate(lost_d))
    d.close()
