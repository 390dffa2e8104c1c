"""The deck runner's host side and the SDDS reader, on the CPU: files written
and read by the pure-Python SDDS checker, the reference's example waveform,
and deck errors that stop before any device work."""
import os

import numpy as np
import pytest

from oracle import sdds_ref
from spiffe_amd import abi, deck

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DECKS = os.path.join(os.path.dirname(HERE), "spiffe_amd", "decks")


def test_ascii_waveform_reads_alike():
    path = os.path.join(GOLD, "spline4.wf")      # examples/apsTCRFGun/excite/spline4.wf
    ref = sdds_ref.read(path)
    assert deck.sdds_info(path) == (1, len(ref["pages"][0]["columns"]["t"]))
    for c in ("t", "W"):
        assert np.array_equal(deck.sdds_column(path, c), ref["pages"][0]["columns"][c])
    assert deck.sdds_column(path, "t")[1] == 7e-12


def test_binary_pages_of_every_numeric_type(tmp_path):
    path = str(tmp_path / "x.sdds")
    cols = [("a", "double"), ("b", "float"), ("c", "long"), ("d", "short")]
    rng = np.random.default_rng(3)
    pages = []
    for rows in (5, 0, 17):
        pages.append(((rows * 1.5, rows), {"a": rng.standard_normal(rows), "b": rng.standard_normal(rows),
                                            "c": rng.integers(-1000, 1000, rows), "d": rng.integers(-300, 300, rows)}))
    sdds_ref.write_binary(path, cols, pages, params=[("time", "double"), ("n", "long")])
    assert deck.sdds_info(path) == (3, 5)
    for k, (pv, cv) in enumerate(pages):
        assert deck.sdds_parameter(path, "time", k) == pv[0]
        assert deck.sdds_parameter(path, "n", k) == pv[1]
        assert np.array_equal(deck.sdds_column(path, "a", k), cv["a"])
        assert np.array_equal(deck.sdds_column(path, "b", k), cv["b"].astype(np.float32).astype(np.float64))
        assert np.array_equal(deck.sdds_column(path, "c", k), cv["c"])
        assert np.array_equal(deck.sdds_column(path, "d", k), cv["d"])


def test_example_profiles_and_waveforms_read():
    t = deck.sdds_column(os.path.join(DECKS, "flat_top_10ps.sdds"), "t")
    assert t.tolist() == [0.0, 0.5e-12, 10.5e-12, 11e-12]
    w = deck.sdds_column(os.path.join(DECKS, "ramp_200dt.wf"), "W")
    assert w.tolist() == [0.0, 1.0, 1.0]


def deck_error(tmp_path, text, files=()):
    for name in files:
        with open(os.path.join(DECKS, name)) as src, open(tmp_path / name, "w") as dst:
            dst.write(src.read())
    p = tmp_path / "t.spiffe"
    p.write_text(text)
    with pytest.raises(deck.DeckError) as e:
        deck.run_deck(str(p))
    return e.value


GEO = '&define_geometry nz=41, nr=21, zmin=0, zmax=0.03, rmax=0.012, boundary="dc_gun.geo" &end\n'


@pytest.mark.parametrize("text,code,msg", [
    ("&define_cathode z_position=0 &end", abi.SPF_EINVAL, "before define_geometry"),
    (GEO + "&define_bogus x=1 &end", abi.SPF_EINVAL, "unknown namelist"),
    (GEO + "&define_emitter material_id=1 &end", abi.SPF_EUNSUPPORTED, "not supported"),
    (GEO.replace("nz=41", "nz=1"), abi.SPF_EINVAL, "nz and nr must be greater than 2"),
    (GEO.replace('boundary="dc_gun.geo"', 'boundary="dc_gun.geo", lower="Neumann"'), abi.SPF_EINVAL, "meaningless"),
    (GEO + "&define_cathode z_position=0, outer_radius=1e-3, colour=2 &end", abi.SPF_EINVAL, "unknown parameter"),
    (GEO + '&define_screen z_position=0.01 &end', abi.SPF_EINVAL, "no filename given for screen"),
    (GEO + "&poisson_correction accuracy=0 &end", abi.SPF_EINVAL, "accuracy <= 0"),
    (GEO + "&integrate dt_integration=0 &end", abi.SPF_EINVAL, "dt_integration <= 0"),
    (GEO + "&integrate dt_integration=1e-9, finish_time=1e-8 &end", abi.SPF_EINVAL, "Courant"),
    (GEO + "&integrate dt_integration=1e-13, finish_time=1e-12, smoothing_parameter=0.5 &end",
     abi.SPF_EUNSUPPORTED, "smoothing"),
    ('&define_geometry nz=41, nr=21, zmin=0, zmax=0.03, rmax=0.012, boundary="nope.geo" &end', abi.SPF_EINVAL,
     "cannot open boundary file"),
])
def test_deck_errors_stop_before_device_work(tmp_path, text, code, msg):
    e = deck_error(tmp_path, text, files=["dc_gun.geo"])
    assert e.code == code and msg in str(e), str(e)
