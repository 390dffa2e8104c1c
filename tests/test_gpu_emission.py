"""Cathode emission through the C ABI vs the oracle, bit for bit.

The library makes the count and the random draws on the host and writes the
new particles on the device (momenta, spread-over-dt drift and kick with the
fields in HBM); the oracle restates src/cathode.c whole.  Both use the same
Philox stream, so every particle column must match exactly.
"""
import math

import numpy as np
import pytest

from _support import bits, pillbox_flags, pillbox_grid, random_fields
from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd.device import SpfError
from spiffe_amd.moments import relative_differences, screen_moments
from test_gpu_parity import assert_fields, assert_particles, assert_rows, ctx, load_fields, pair  # noqa: F401

pytestmark = pytest.mark.gpu

TMI = abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP


def gaussian_profile(t0, sigma, n=81):
    t = t0 + sigma * np.linspace(-4, 4, n)
    return t, np.exp(-0.5 * ((t - t0) / sigma) ** 2)


def set_step(o, d, step):
    o.f.time_step = step
    o.f.time = step * o.f.dti + o.f.start_time
    d.set_step(step)


def emit_both(o, d, ks, steps, start=0):
    counts = []
    for step in range(start, start + steps):
        set_step(o, d, step)
        for k in ks:
            no = o.emit_electrons_from_cathode(k)
            nd = d.emit_electrons_from_cathode(k)
            assert no == nd, "cathode %d step %d: oracle emitted %d, device %d" % (k, step, no, nd)
            counts.append(no)
    return counts


def assert_state(o, d, k):
    so, sd = o.cathode_state(k), d.cathode_state(k)
    for n in ["current_density", "ref_current_density", "number_per_step", "electrons_per_macroparticle", "Q",
              "temperature", "sigma_v", "initial_vz", "start_step", "stop_step", "n_bins"]:
        assert bits(np.float64(so[n])) == bits(np.float64(sd[n])), "cathode state %s: %r vs %r" % (n, so[n], sd[n])


CASES = {
    # uniform in area with random rounding, thermal kicks, Gaussian time profile
    "unzoned_thermal_profile": dict(
        steps=40, cap=60000,
        kw=dict(z_position=0.0, outer_radius=1e-3, current_density=2.0e6, electrons_per_macroparticle=2.0e5,
                zoned_emission=0, temperature=5802.0, add_thermal_velocities=1, initial_pz=1.978e-3,
                random_number_seed=987654321),
        profile=lambda dt: gaussian_profile(20 * dt, 5 * dt)),
    # per-annulus random rounding (the default)
    "zoned": dict(steps=30, cap=40000,
                  kw=dict(z_position=0.0, outer_radius=2.2e-3, current_density=1.0e6,
                          electrons_per_macroparticle=1.0e5, initial_pz=1e-3, random_number_seed=77)),
    # distribution correction every 3 steps, annuli off the axis
    "correction": dict(steps=30, cap=20000,
                       kw=dict(z_position=0.0, inner_radius=1.1e-3, outer_radius=3.0e-3, current_density=2.0e4,
                               electrons_per_macroparticle=1.0e4, distribution_correction_interval=3,
                               initial_pz=2e-3, random_number_seed=5)),
    # equal-area radii with the spread-over-dt kick through TM fields in HBM
    "discretize_spread_tm": dict(steps=12, cap=20000, fields=True,
                                 kw=dict(z_position=0.0, inner_radius=3e-4, outer_radius=1.7e-3,
                                         current_density=1.0e6, number_per_step=700.0, discretize_radii=1,
                                         zoned_emission=0, spread_over_dt=1, initial_pz=5e-3,
                                         initial_omega=1e9, temperature=3000.0, add_thermal_velocities=1)),
    # Halton radii and emission times with the kick
    "halton_spread": dict(steps=15, cap=20000, fields=True,
                          kw=dict(z_position=0.0, outer_radius=1.3e-3, current_density=1.0e6,
                                  number_per_step=333.3, zoned_emission=0, halton_radix_r=3, halton_radix_dt=2,
                                  spread_over_dt=1, initial_pz=3e-3)),
    # window and autophase: emission starts time_offset after the first call
    "autophase_window": dict(steps=40, cap=20000,
                             kw=dict(z_position=0.0, outer_radius=1e-3, current_density=1.0e6,
                                     number_per_step=97.7, autophase=1, start_time=0.0, stop_time=12e-12,
                                     time_offset=5e-12, zoned_emission=0, random_number_seed=31)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_emission_modes_bitwise(ctx, case):
    spec = CASES[case]
    g = pillbox_grid(modes=TMI)
    f = pillbox_flags(g)
    o, d = pair(ctx, g, f, cap=spec["cap"])
    if spec.get("fields"):
        fl = random_fields(g, 17, names=["Ez", "Er", "EzImposed", "ErImposed"], scale=3e7)
        load_fields(o, d, fl)
    profile = spec["profile"](g.dt) if "profile" in spec else None
    ko = o.define_cathode(profile=profile, **spec["kw"])
    kd = d.define_cathode(profile=profile, **spec["kw"])
    assert ko == kd == [0]
    counts = emit_both(o, d, ko, spec["steps"])
    assert sum(counts) > 100, "case %s emitted only %d particles" % (case, sum(counts))
    assert_state(o, d, 0)
    assert_particles(o, d)
    assert d.count() == (sum(counts), sum(counts))
    assert d.next_particle_id() == 1 + sum(counts)
    d.close()


def test_richardson_schottky_and_field_emission_follow_the_field(ctx):
    g = pillbox_grid(modes=TMI)
    f = pillbox_flags(g)
    o, d = pair(ctx, g, f, cap=50000)
    # an accelerating field for electrons (Ez < 0) that varies from point to point
    fl = random_fields(g, 23, names=["Ez", "EzImposed"], scale=1e6)
    fl["Ez"] -= 3e7
    load_fields(o, d, fl)
    rs = dict(z_position=0.0, outer_radius=1.5e-3, current_density=1e5, work_function=2.0, determine_temperature=1,
              electrons_per_macroparticle=2e4, initial_pz=1e-3, zoned_emission=0, random_number_seed=3)
    fe = dict(z_position=0.0, inner_radius=0.0, outer_radius=2.3e-3, field_emission=1, work_function=4.5,
              field_emission_beta=80.0, electrons_per_macroparticle=1e2, initial_pz=1e-3)
    ks = o.define_cathode(**rs) + o.define_cathode(**fe)
    assert d.define_cathode(**rs) + d.define_cathode(**fe) == ks
    assert len(ks) == 1 + int((2.3e-3 + g.dr / 2) / g.dr)
    for k in ks:
        o.set_current_density(k)
        d.set_current_density(k)
        assert_state(o, d, k)
    emit_both(o, d, ks, 6)
    for k in ks:
        assert_state(o, d, k)
    assert_particles(o, d)
    d.close()


def test_capacity_is_checked_before_emission(ctx):
    g = pillbox_grid(modes=TMI)
    f = pillbox_flags(g)
    o, d = pair(ctx, g, f, cap=100)
    (k,) = d.define_cathode(z_position=0.0, outer_radius=1e-3, current_density=1.0, number_per_step=150.0,
                            zoned_emission=0)
    d.set_step(0)
    with pytest.raises(SpfError) as e:
        d.emit_electrons_from_cathode(k)
    assert e.value.code == abi.SPF_ECAPACITY
    assert d.count() == (0, 0)
    d.close()


def test_pillbox_emission_and_push(ctx):
    """The small plain-method pattern: a pillbox in its TM010 mode, a 1 mm cathode on the
    z=0 wall emitting a Gaussian pulse (sigma 2 ps, ~10 k macroparticles of 0.1 nC in
    all, 1 eV), space charge on, a screen; emission runs inside each step.  Over the
    first 1000 steps the emitted, lost and screen-crossing counts agree step by step,
    the particles (hence their cell indices) bit for bit, and the bunch moments at
    the screen within 1e-6 relative."""
    from scipy.special import j0
    g = pillbox_grid(modes=TMI | abi.SPACE_CHARGE)
    f = pillbox_flags(g)
    screens = [(0.005, 1)]
    o, d = pair(ctx, g, f, cap=16000, screens=screens, compat=False)
    # TM010: Ez = E0 J0(2.405 r/R), 20 MV/m, starting in the phase that accelerates electrons
    r = g.dr * np.arange(g.nr + 1)
    Ez = np.tile(-2.0e7 * j0(2.404825557695773 * r / g.rmax), (g.nz + 1, 1))
    Ez[:, g.nr - 1:] = 0
    load_fields(o, d, {"Ez": Ez})
    sigma = 2e-12
    Q = 0.1e-9
    epm = Q / K.e_mks / 10000
    I_peak = Q / (sigma * math.sqrt(2 * math.pi))
    J = I_peak / (math.pi * 1e-6)
    kw = dict(z_position=0.0, outer_radius=1e-3, current_density=J, electrons_per_macroparticle=epm,
              initial_pz=1.978e-3, zoned_emission=0, random_number_seed=987654321)
    prof = gaussian_profile(4 * sigma, sigma)
    assert o.define_cathode(profile=prof, **kw) == d.define_cathode(profile=prof, **kw) == [0]
    rows_o, rows_d, lost_o, lost_d = [], [], [], []
    for step in range(1000):
        ro, lo = o.step(screens=screens, compat=False)
        d.step()
        rd, ld = d.drain()
        rows_o.append(ro), rows_d.append(rd), lost_o.append(lo), lost_d.append(ld)
        assert len(ro) == len(rd) and len(lo) == len(ld), "event counts differ at step %d" % step
    n, _ = d.count()
    emitted = d.next_particle_id() - 1
    assert 9000 < emitted < 11000, emitted
    # by step 1000 the pulse has crossed the 2.5 cm cavity: every particle that is
    # not left is on the lost list
    assert n + sum(len(x) for x in lost_d) == emitted
    assert_fields(o, d, ["Ez", "Er", "Bphi", "Jz", "Jr"])
    assert_particles(o, d)
    ro, rd = np.concatenate(rows_o), np.concatenate(rows_d)
    assert_rows(ro, rd, np.concatenate(lost_o), np.concatenate(lost_d))
    assert len(rd) > 1000, "too few screen crossings (%d) for the moment check" % len(rd)
    diffs = relative_differences(screen_moments(ro), screen_moments(rd))
    assert max(diffs.values()) <= 1e-6, diffs
    d.close()
