"""Bunch moments at a screen, from the rows a run drains.

A screen row (src/screens.c:82-98, filled at src/advance_particles.c:371-420)
holds the crossing radius r, the normalised momenta pz, pr, pphi (beta*gamma
components), the crossing time t and the macroparticle charge q.  The
reference writes these rows to SDDS and leaves the moments to post-processing;
this module computes the four the design scan reports:

- charge: the sum of q over the rows;
- mean gamma: the |q|-weighted mean of sqrt(1 + pz^2 + pr^2 + pphi^2);
- rms bunch length: the |q|-weighted rms spread of the crossing time, and
  the same times the mean longitudinal velocity <pz/gamma> c as a length;
- normalised rms emittance of one transverse plane of a round beam: with
  x = r cos(phi), px = pr cos(phi) - pphi sin(phi) and phi uniform,
  <x^2> = <r^2>/2, <px^2> = (<pr^2> + <pphi^2>)/2, <x px> = <r pr>/2, so
  eps_n,x = sqrt(<x^2><px^2> - <x px>^2)
          = 0.5 sqrt(<r^2>(<pr^2> + <pphi^2>) - <r pr>^2)   [m].

All sums are plain numpy float64 reductions of the rows in the order given.
"""
import numpy as np

C_MKS = 2.99792458e8


def screen_moments(rows, screen=None):
    """Moments of the screen rows (a structured array with fields r, pz, pr,
    pphi, t, q and screen), optionally of one screen only."""
    if screen is not None:
        rows = rows[rows["screen"] == screen]
    n = len(rows)
    if n == 0:
        return {"n": 0, "charge": 0.0, "mean_gamma": float("nan"), "sigma_t": float("nan"),
                "rms_length": float("nan"), "emittance_n": float("nan")}
    r = np.asarray(rows["r"], dtype=np.float64)
    pz = np.asarray(rows["pz"], dtype=np.float64)
    pr = np.asarray(rows["pr"], dtype=np.float64)
    pphi = np.asarray(rows["pphi"], dtype=np.float64)
    t = np.asarray(rows["t"], dtype=np.float64)
    q = np.asarray(rows["q"], dtype=np.float64)
    w = np.abs(q)
    wsum = w.sum()
    gamma = np.sqrt(1.0 + pz * pz + pr * pr + pphi * pphi)

    def mean(x):
        return float((w * x).sum() / wsum)

    tbar = mean(t)
    sigma_t = np.sqrt(mean((t - tbar) ** 2))
    beta_z = mean(pz / gamma)
    r2, pr2, pphi2, rpr = mean(r * r), mean(pr * pr), mean(pphi * pphi), mean(r * pr)
    eps = 0.5 * np.sqrt(max(r2 * (pr2 + pphi2) - rpr * rpr, 0.0))
    return {"n": n, "charge": float(q.sum()), "mean_gamma": mean(gamma), "sigma_t": float(sigma_t),
            "rms_length": float(sigma_t * beta_z * C_MKS), "emittance_n": float(eps)}


def relative_differences(a, b):
    """Largest relative difference of each moment between two results."""
    out = {}
    for k in ("charge", "mean_gamma", "sigma_t", "rms_length", "emittance_n"):
        x, y = a[k], b[k]
        scale = max(abs(x), abs(y))
        out[k] = 0.0 if scale == 0 else abs(x - y) / scale
    return out
