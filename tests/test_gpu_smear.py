"""Area-weighted ("smeared") deposition and the Poisson guesses through the C
ABI vs the oracle, bit for bit.

Smearing (src/space_charge.c and src/poisson_correction.c:188-239 with
FL_RADIAL_SMEAR / FL_LONGIT_SMEAR) gives each particle up to four target
points per component; the device keeps the reference's per-point summation
order (particle index order), zero terms included.  The line- and
point-charge guesses (src/poisson_correction.c:564-615) start the solve from
Qerror and the charge moment summed in the reference's serial order.
"""
import numpy as np
import pytest

from _support import TMTE, gun_flags, gun_grid, random_fields, random_particles
from oracle.oracle import OracleSim
from spiffe_amd import abi
from spiffe_amd.device import DeviceRun
from test_gpu_parity import _edge_particles, assert_fields, assert_particles, ctx, load_fields, pair  # noqa: F401

pytestmark = pytest.mark.gpu

J_NAMES = ["Jz", "Jr", "Jphi"]
SMEARS = {"radial": abi.RADIAL_SMEAR, "longitudinal": abi.LONGIT_SMEAR,
          "both": abi.RADIAL_SMEAR | abi.LONGIT_SMEAR}


@pytest.mark.parametrize("smear", sorted(SMEARS))
@pytest.mark.parametrize("n,seed", [(1, 1), (777, 2), (20000, 3)])
def test_smeared_deposit_bitwise(ctx, smear, n, seed):
    """Particles on grid lines and half-lines, on the axis, in the top rows and
    past both z ends (where the reference writes row nz+1, past its arrays,
    and both sides leave that term out), with inactive ones among them."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE | SMEARS[smear])
    f = gun_flags(g)
    o, d = pair(ctx, g, f, cap=n)
    p = _edge_particles(g, n, seed) if n > 100 else random_particles(g, n, seed)
    p["status"][::7] = 0
    o.append(p)
    d.append(p)
    o.add_particle_currents()
    d.add_particle_currents()
    assert_fields(o, d, J_NAMES)
    d.close()


def test_smeared_deposit_concentrated_bunch(ctx):
    """Most particles in a few cells: point segments thousands of terms long."""
    n = 300_001
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE | abi.RADIAL_SMEAR | abi.LONGIT_SMEAR)
    f = gun_flags(g)
    o, d = pair(ctx, g, f, cap=n)
    p = random_particles(g, n, 5)
    rng = np.random.default_rng(105)
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


@pytest.mark.parametrize("smear", sorted(SMEARS))
@pytest.mark.parametrize("thr", [0.0, -1e-9])
def test_smeared_poisson_correction_steps(ctx, smear, thr):
    """Whole steps with smeared deposition and Poisson corrections every other
    step: the smeared charge removal, Qreal (a relative error-charge
    threshold) and the corrected fields."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE | SMEARS[smear])
    f = gun_flags(g)
    n = 3000
    o, d = pair(ctx, g, f, cap=n)
    load_fields(o, d, random_fields(g, 41, names=["Ez", "Er", "Bphi"], scale=1e5))
    p = random_particles(g, n, 42, flags=f, q=-1e-13)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=2, accuracy=1e-9, maximum_iterations=4000,
                                     error_charge_threshold=thr)
    its_o, its_d = [], []
    for step in range(4):
        o.step(screens=(), compat=True)
        d.step()
        its_o.append(o.pc.last_iterations)
        its_d.append(d.poisson_stats()[0])
    assert any(i > 0 for i in its_o)
    assert all(a == b for a, b in zip(its_o, its_d) if a > 0), (its_o, its_d)
    assert_fields(o, d, ["Ez", "Er", "Bphi", "Psi", "Q", "Jz", "Jr"])
    assert_particles(o, d)
    d.close()


def test_cell_sort_run_deposits_smeared_in_its_particle_order(ctx):
    """A cell-sort run with smearing takes the ordered deposition: after its
    re-sort, its steps match the oracle run from the sorted particles."""
    g = gun_grid(modes=TMTE | abi.SPACE_CHARGE | abi.RADIAL_SMEAR | abi.LONGIT_SMEAR)
    f = gun_flags(g)
    n = 20000
    d = DeviceRun(ctx, g, f, max_particles=n, compat=False, cell_sort=5)
    p = random_particles(g, n, 17, flags=f)
    d.append(p)
    d.sort_particles()
    sorted_p = d.particles()
    assert not np.array_equal(sorted_p["particleID"], p["particleID"]), "the sort should reorder the particles"
    o = OracleSim(g, f, max_particles=n)
    fl = random_fields(g, 5)
    load_fields(o, d, fl)
    o.append(sorted_p)
    for _ in range(3):
        o.step(screens=(), compat=False)
        d.step()
    assert_fields(o, d, ["Ez", "Er", "Bphi", "Ephi", "Br", "Bz"] + J_NAMES)
    assert_particles(o, d)
    d.close()


@pytest.mark.parametrize("code", [1, 2])
def test_poisson_guess_steps(ctx, code):
    """Corrections that start from the line-charge (1) or point-charge (2)
    guess: Psi, the iteration counts and the corrected fields."""
    g = gun_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE)
    f = gun_flags(g)
    n = 3000
    o, d = pair(ctx, g, f, cap=n)
    load_fields(o, d, random_fields(g, 43, names=["Ez", "Er", "Bphi"], scale=1e5))
    p = random_particles(g, n, 44, flags=f, q=-1e-13)
    o.append(p)
    d.append(p)
    for obj in (o, d):
        obj.setup_poisson_correction(step_interval=2, accuracy=1e-9, maximum_iterations=4000, guess_code=code)
    its_o, its_d = [], []
    for step in range(4):
        o.step(screens=(), compat=True)
        d.step()
        its_o.append(o.pc.last_iterations)
        its_d.append(d.poisson_stats()[0])
    assert its_o[0] > 0 and its_o[2] > 0
    assert (its_o[0], its_o[2]) == (its_d[0], its_d[2]), (its_o, its_d)
    assert_fields(o, d, ["Ez", "Er", "Bphi", "Psi", "Q"])
    assert_particles(o, d)
    d.close()
