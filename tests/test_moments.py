"""Screen moments (spiffe_amd.moments) pinned by closed forms and by a
brute-force Cartesian evaluation of the round-beam emittance."""
import numpy as np

from spiffe_amd.device import SCREEN_ROW_DTYPE
from spiffe_amd.moments import relative_differences, screen_moments


def rows_of(r, pz, pr, pphi, t, q, screen=0):
    rows = np.zeros(len(r), dtype=SCREEN_ROW_DTYPE)
    rows["r"], rows["pz"], rows["pr"], rows["pphi"], rows["t"], rows["q"] = r, pz, pr, pphi, t, q
    rows["screen"] = screen
    return rows


def random_rows(n, seed):
    rng = np.random.default_rng(seed)
    return rows_of(1e-3 * np.sqrt(rng.random(n)), 2.0 + 0.05 * rng.standard_normal(n),
                   0.01 * rng.standard_normal(n), 0.003 * rng.standard_normal(n),
                   1e-9 + 3e-12 * rng.standard_normal(n), np.full(n, -1e-15) * (1 + rng.random(n)))


def test_emittance_equals_cartesian_expansion():
    """Each row expanded into four Cartesian particles at phi = 0, pi/2, pi, 3pi/2
    (x = r cos phi, px = pr cos phi - pphi sin phi) has <x> = <px> = 0 and the
    round-beam averages exactly, so the plain 2-D emittance of the expansion is
    the moment's value."""
    rows = random_rows(5000, 3)
    m = screen_moments(rows)
    r, pr, pphi, w = rows["r"], rows["pr"], rows["pphi"], np.abs(rows["q"])
    xs, pxs, ws = [], [], []
    for c, s in ((1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)):
        xs.append(r * c)
        pxs.append(pr * c - pphi * s)
        ws.append(w)
    x, px, ww = np.concatenate(xs), np.concatenate(pxs), np.concatenate(ws)
    W = ww.sum()
    x2, px2, xpx = (ww * x * x).sum() / W, (ww * px * px).sum() / W, (ww * x * px).sum() / W
    eps = np.sqrt(x2 * px2 - xpx * xpx)
    assert abs(m["emittance_n"] - eps) <= 1e-12 * eps


def test_closed_forms():
    n = 1000
    r = np.linspace(1e-5, 1e-3, n)
    # laminar: pr proportional to r, no pphi -> zero emittance
    m = screen_moments(rows_of(r, np.full(n, 2.0), 0.3 * r, np.zeros(n), np.zeros(n), np.full(n, -2e-15)))
    assert m["emittance_n"] <= 1e-12 * np.sqrt(np.mean(r * r) * np.mean((0.3 * r) ** 2))
    # monoenergetic along z: gamma = sqrt(1 + pz^2), charge = sum of q
    m = screen_moments(rows_of(r, np.full(n, 2.0), np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, -2e-15)))
    assert abs(m["mean_gamma"] - np.sqrt(5.0)) <= 4e-15 * np.sqrt(5.0)
    assert abs(m["charge"] + 2e-15 * n) <= 1e-12 * 2e-15 * n
    # two equal charges at t0 -/+ d: sigma_t = d, length = d * beta c
    d = 2e-12
    m = screen_moments(rows_of(np.zeros(2), np.full(2, 2.0), np.zeros(2), np.zeros(2), np.array([1e-9 - d, 1e-9 + d]),
                               np.full(2, -1e-15)))
    assert abs(m["sigma_t"] - d) <= 1e-6 * d
    beta = 2.0 / np.sqrt(5.0)
    assert abs(m["rms_length"] - d * beta * 2.99792458e8) <= 1e-6 * m["rms_length"]
    # uncorrelated, equal weights, pphi only: <r^2> = a, <pphi^2> = b -> eps = 0.5 sqrt(a b)
    rr = np.array([1e-3, 2e-3, 1e-3, 2e-3])
    pp = np.array([0.01, 0.01, -0.01, -0.01])
    m = screen_moments(rows_of(rr, np.full(4, 2.0), np.zeros(4), pp, np.zeros(4), np.full(4, -1e-15)))
    a, b = np.mean(rr * rr), 1e-4
    assert abs(m["emittance_n"] - 0.5 * np.sqrt(a * b)) <= 1e-12 * m["emittance_n"]


def test_screen_selection_and_differences():
    a = random_rows(300, 5)
    b = random_rows(200, 6)
    b["screen"] = 1
    both = np.concatenate([a, b])
    assert screen_moments(both, screen=1) == screen_moments(b)
    assert screen_moments(both, screen=2)["n"] == 0
    d = relative_differences(screen_moments(a), screen_moments(a))
    assert max(d.values()) == 0.0
    a2 = a.copy()
    a2["pz"] *= 1 + 1e-9
    d = relative_differences(screen_moments(a), screen_moments(a2))
    assert 0 < d["mean_gamma"] < 1e-8
