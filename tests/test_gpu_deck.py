"""Decks run end to end by the library's deck runner (spf_deck_run), checked
against the same run assembled from the oracle: every screen page, lost row
and saved field page bit for bit (files read with the pure-Python SDDS
checker)."""
import math
import os
import shutil

import numpy as np
import pytest

from _support import bits
from oracle import geometry_ref as G
from oracle import oracle as O
from oracle import sdds_ref
from spiffe_amd import abi, deck
from spiffe_amd import constants as K
from spiffe_amd.device import Grid

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(os.path.dirname(HERE), "spiffe_amd", "decks")
SCREEN_COLS = ["r", "pz", "pr", "pphi", "t", "q", "r0", "t0"]


def oracle_flags(nz, nr, zmin, zmax, rmax, geo):
    with open(os.path.join(DECKS, geo)) as fp:
        L, P, _ = G.draw(nz, nr, zmin, zmax, rmax, G.parse_points(fp.read()))
    return O.derive_flags(np.array(L, dtype=np.int16)), np.array(P)


def profile(name):
    f = sdds_ref.read(os.path.join(DECKS, name))["pages"][0]["columns"]
    return f["t"], f["A"]


def compare_screen(path, screen_rows):
    """screen_rows: list over steps of (step, rows of this screen)."""
    pages = sdds_ref.read(path)["pages"]
    want = [(s, r) for s, r in screen_rows if len(r)]
    assert len(pages) == len(want), "%s: %d pages, oracle %d" % (path, len(pages), len(want))
    for pg, (step, r) in zip(pages, want):
        assert pg["params"]["Timestep"] == step
        c = pg["columns"]
        for n in SCREEN_COLS:
            assert np.array_equal(bits(c[n]), bits(r[n])), "%s step %d column %s" % (path, step, n)
        for a, b in (("rMin", "rmin"), ("rMax", "rmax"), ("zMin", "zmin"), ("zMax", "zmax")):
            assert np.array_equal(bits(c[a]), bits(r[b]))
        assert np.array_equal(c["particleID"], r["particleID"]) and np.array_equal(c["generation"], r["generation"])
    return sum(len(r) for _, r in want)


def compare_lost(path, lost_rows):
    pages = sdds_ref.read(path)["pages"]
    want = [l for l in lost_rows if len(l)]
    assert len(pages) == len(want)
    for pg, l in zip(pages, want):
        c = pg["columns"]
        assert np.array_equal(c["t"], l["t"])
        for n in ("rLost", "zLost", "pz", "pr", "z0", "r0", "q"):
            assert np.array_equal(c[n], l[n].astype(np.float32)), n
        assert np.array_equal(c["particleID"], l["particleID"])
    return sum(len(l) for l in want)


@pytest.mark.parametrize("nz,nr,N,dt,steps,interval,options", [
    (151, 61, 20000, 4e-13, 700, 10, False),        # 0.2 mm mesh, 20 k macroparticles, 700 steps
    (151, 61, 20000, 4e-13, 700, 10, True),         # the same with the options below
    (601, 241, 200000, 1e-13, 1600, 100, False),    # the serving stream's DC gun: 0.05 mm, 200 k, its own dt
])
def test_dc_gun_deck(ctx_free, tmp_path, nz, nr, N, dt, steps, interval, options):
    """The DC-gun deck with field saving; on the full 0.05 mm mesh with all 200 k
    macroparticles over the first 1600 steps (the bunch reaches the first screen),
    with a Poisson correction every 100 steps so the restatement keeps up.  With
    `options` the deck also sets radial and longitudinal smearing
    (define_geometry), the point-charge guess (poisson_correction) and
    J_filter_multiplier (integrate)."""
    for f in ("dc_gun.geo", "flat_top_10ps.sdds"):
        shutil.copy(os.path.join(DECKS, f), tmp_path / f)
    text = open(os.path.join(DECKS, "dc_gun.spiffe")).read()
    epm = 50e-12 / K.e_mks / N
    text = text.replace("nz = 601, nr = 241", "nz = %d, nr = %d" % (nz, nr))
    text = text.replace("electrons_per_macroparticle = 1560.5454", "electrons_per_macroparticle = %r" % epm)
    text = text.replace("step_interval = 10,", "step_interval = %d," % interval)
    text = text.replace("dt_integration = 1e-13, finish_time = 400e-12",
                        "dt_integration = %r, finish_time = %r" % (dt, (steps + 0.5) * dt))
    text = text.replace("&integrate", '&define_field_saving filename = "%%s.fsave", time_interval = %r, '
                                      'save_before_exiting = 1 &end\n\n&integrate' % (100 * dt))
    if options:
        text = text.replace('boundary = "dc_gun.geo"',
                            'boundary = "dc_gun.geo", radial_smearing = 1, longitudinal_smearing = 1')
        text = text.replace("maximum_iterations = 2000", 'maximum_iterations = 2000, guess_type = "point-charge"')
        text = text.replace("space_charge = 1,", "space_charge = 1, J_filter_multiplier = 0.2,")
        assert text.count("smearing") == 2 and "point-charge" in text and "J_filter" in text
    (tmp_path / "dc.spiffe").write_text(text)
    cap_particles = int(1.5 * N)
    stats = deck.run_deck(str(tmp_path / "dc.spiffe"), max_particles=cap_particles, compat=False)
    assert stats["steps"] == steps

    # the same run from the oracle
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=0.030, rmax=0.012, dt=dt,
             modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE |
             (abi.RADIAL_SMEAR | abi.LONGIT_SMEAR if options else 0))
    f, psi = oracle_flags(nz, nr, 0.0, 0.030, 0.012, "dc_gun.geo")
    o = O.OracleSim(g, f, max_particles=cap_particles)
    o.set_field("PsiImposed", psi)
    o.setup_poisson_correction(step_interval=interval, accuracy=1e-6, maximum_iterations=2000,
                               guess_code=2 if options else 0)
    o.define_cathode(profile=profile("flat_top_10ps.sdds"), z_position=0.0, outer_radius=3e-3,
                     current_density=1.768388e5, electrons_per_macroparticle=epm, initial_pz=1e-3,
                     zoned_emission=0, random_number_seed=987654321)
    o.compute_imposed_Efield(True)
    o.f.time_step, o.f.time = 0, 0.0
    o.add_poisson_correction(True)
    o.set_current_density(0)
    screens = [(0.010, 1), (0.025, 1)]
    rows_by_screen = [[], []]
    lost_pages, saves = [], []
    for step in range(steps):
        if step % 100 == 0:
            saves.append((step, {n: o.field(n) for n in ("Ez", "Er", "Jz", "Jr", "Bphi", "Bphi2")}))
        rows, lost = o.step(screens=screens, compat=False, j_filter=0.2 if options else 0.0)
        for s in range(2):
            rows_by_screen[s].append((step, rows[rows["screen"] == s]))
        lost_pages.append(lost)
    saves.append((steps, {n: o.field(n) for n in ("Ez", "Er", "Jz", "Jr", "Bphi", "Bphi2")}))
    n1 = compare_screen(str(tmp_path / "dc.sc1"), rows_by_screen[0])
    n2 = compare_screen(str(tmp_path / "dc.sc2"), rows_by_screen[1])
    nl = compare_lost(str(tmp_path / "dc.lost"), lost_pages)
    if N == 20000:
        assert n1 > 0.9 * N and n2 > 0 and nl > 0
    else:
        assert n1 > 0
    assert stats["screen_rows"] == n1 + n2 and stats["lost_rows"] == nl
    fs = sdds_ref.read(str(tmp_path / "dc.fsave"))["pages"]
    assert len(fs) == len(saves) == stats["field_saves"]
    for pg, (step, fl) in zip(fs, saves):
        assert pg["params"]["time"] == step * dt
        for n, a in fl.items():
            assert np.array_equal(bits(pg["columns"][n]), bits(a.ravel())), "saved %s at step %d" % (n, step)
    assert np.array_equal(fs[0]["columns"]["metalFlags"], f.ravel().astype(np.int32))


def test_pillbox_deck_antenna_and_emission(ctx_free, tmp_path):
    # the plain-method deck whole, 5200 steps: the antenna rings the cavity up
    # (200-step ramp, 2000 field-only steps), the cathode emits from step 2200,
    # 3000 steps with the beam; every screen page and lost row bit for bit
    for fn in ("pillbox.geo", "ramp_200dt.wf", "gauss_2ps.sdds"):
        shutil.copy(os.path.join(DECKS, fn), tmp_path / fn)
    shutil.copy(os.path.join(DECKS, "pillbox.spiffe"), tmp_path / "pb.spiffe")
    steps = 5200
    stats = deck.run_deck(str(tmp_path / "pb.spiffe"), max_particles=16000, max_steps=steps)
    assert stats["steps"] == steps and stats["emitted"] > 1000
    dt = 1.061394451537358e-12
    g = Grid(nz=51, nr=81, zmin=0.0, zmax=0.025, rmax=0.04, dt=dt,
             modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP)
    f, _ = oracle_flags(51, 81, 0.0, 0.025, 0.04, "pillbox.geo")
    o = O.OracleSim(g, f, max_particles=16000)
    o.define_cathode(profile=profile("gauss_2ps.sdds"), z_position=0.0, outer_radius=1e-3,
                     current_density=6.3493e6, electrons_per_macroparticle=62415.09, start_time=2.335068e-09,
                     initial_pz=1.978e-3, zoned_emission=0, random_number_seed=987654321)
    wf = sdds_ref.read(os.path.join(DECKS, "ramp_200dt.wf"))["pages"][0]["columns"]
    t_wf, W = np.ascontiguousarray(wf["t"]), np.ascontiguousarray(wf["W"])
    import ctypes
    lib = O.load()
    last = ctypes.c_int64(0)
    errc = ctypes.c_int32(0)
    iend = int(0.025 / g.dz + 0.5)
    rows_by_step, lost_pages = [], []
    for step in range(steps):
        t = step * dt + dt / 2
        J = lib.oracle_antenna_current(t_wf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       W.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(t_wf),
                                       ctypes.byref(last), 1e9, 2.8685e9, 90.0, t, ctypes.byref(errc))
        rows, lost = o.step(lines=[(0, 0, 0, iend, J)], screens=[(0.012, 1)], compat=True)
        rows_by_step.append((step, rows))
        lost_pages.append(lost)
    n_screen = compare_screen(str(tmp_path / "pb.sc1"), rows_by_step)
    n_lost = compare_lost(str(tmp_path / "pb.lost"), lost_pages)
    assert stats["emitted"] == lib.oracle_cathodes_next_id(o.cathodes) - 1
    assert n_screen > 9000, (n_screen, n_lost, stats)   # the bunch crosses the screen


@pytest.fixture
def ctx_free():
    """The deck runner makes its own context; nothing to share."""
    yield None


def test_saved_fields_reload_scaled_to_ez_peak(ctx_free, tmp_path):
    """excite -> save -> load_fields with Ez_peak: the loaded fields are the saved
    ones times Ez_peak / (extremal on-axis Ez), added onto zero fields."""
    for fn in ("pillbox.geo", "ramp_200dt.wf"):
        shutil.copy(os.path.join(DECKS, fn), tmp_path / fn)
    geo = ('&define_geometry nz = 51, nr = 81, zmin = 0.0, zmax = 0.025, rmax = 0.04, boundary = "pillbox.geo", '
           'include_TE_fields = 1 &end\n')
    ant = ('&define_antenna start = 0.0, end = 0.025, position = 0.0, direction = "z", current = 1.0, '
           'frequency = 2.8685e9, waveform = "ramp_200dt.wf" &end\n')
    dt = 1.061394451537358e-12
    (tmp_path / "excite.spiffe").write_text(
        geo + ant + '&define_field_saving filename = "%%s.fsave", time_interval = 1.0, start_time = 1.0, '
                    'save_before_exiting = 1 &end\n'
        '&integrate dt_integration = %r, finish_time = %r &end\n' % (dt, 400.5 * dt))
    st = deck.run_deck(str(tmp_path / "excite.spiffe"))
    assert st["field_saves"] == 1
    (tmp_path / "load.spiffe").write_text(
        geo + '&load_fields filename = "excite.fsave", Ez_peak = 120e6 &end\n'
              '&define_field_saving filename = "%%s.fsave", time_interval = 1.0 &end\n'
              '&integrate dt_integration = %r, finish_time = %r &end\n' % (dt, 1.5 * dt))
    st = deck.run_deck(str(tmp_path / "load.spiffe"))
    saved = sdds_ref.read(str(tmp_path / "excite.fsave"))["pages"][-1]["columns"]
    loaded = sdds_ref.read(str(tmp_path / "load.fsave"))["pages"][0]["columns"]
    nr1 = 82
    ez_axis = saved["Ez"][::nr1]
    ez_max = 0.0
    for v in ez_axis[1:]:                    # src/load_fields.c:131-145
        if abs(ez_max) < abs(v):
            ez_max = v
    factor = 120e6 / ez_max
    for n in ("Ez", "Er", "Jz", "Jr", "Bphi", "Bphi2", "Ephi", "Jphi", "Br", "Br2", "Bz", "Bz2"):
        assert np.array_equal(bits(loaded[n]), bits(0.0 + saved[n] * factor)), n
    assert np.max(np.abs(loaded["Ez"][::nr1][1:])) == pytest.approx(120e6, rel=1e-15)
