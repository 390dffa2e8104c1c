"""Shared scenario builders for the tests (seeded, small enough for the oracle)."""
import numpy as np

from spiffe_amd import abi
from spiffe_amd import constants as K
from spiffe_amd.device import Grid, courant_dt
from spiffe_amd.geometry import derive_flags, pillbox_marks, polygon_marks

TMTE = abi.TM_FIELDS | abi.TE_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP


def pillbox_grid(modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP, nz=51, nr=81, L=0.025, R=0.04):
    """Closed pillbox of radius R, length L (dz = dr = 0.5 mm for the defaults), dt = 0.9 Courant."""
    dz, dr = L / (nz - 1), R / (nr - 1)
    return Grid(nz=nz, nr=nr, zmin=0.0, zmax=L, rmax=R, dt=courant_dt(dz, dr, 0.9), modes=modes)


def pillbox_flags(grid):
    return derive_flags(pillbox_marks(grid.nz, grid.nr))


def gun_grid(nz=141, nr=61, modes=TMTE):
    """A 1.6-cell-gun-like cavity on a coarse mesh: half cell, iris, full cell, beam pipe."""
    L, R = 0.14, 0.06
    dz, dr = L / (nz - 1), R / (nr - 1)
    return Grid(nz=nz, nr=nr, zmin=0.0, zmax=L, rmax=R, dt=courant_dt(dz, dr, 0.9), modes=modes)


GUN_POLY = [(0.0, 0.0), (0.0, 0.042), (0.030, 0.042), (0.030, 0.0125), (0.035, 0.0125), (0.035, 0.042),
            (0.105, 0.042), (0.105, 0.010), (0.1395, 0.010), (0.1395, 0.0), (0.0, 0.0)]


def gun_flags(grid):
    return derive_flags(polygon_marks(grid.nz, grid.nr, grid.zmin, grid.zmax, grid.rmax, GUN_POLY))


def random_fields(grid, seed, names=("Ez", "Er", "Bphi", "Ephi", "Br", "Bz"), scale=1e6):
    rng = np.random.default_rng(seed)
    out = {}
    for n in names:
        s = scale if n[0] == "E" else scale / K.c_mks
        out[n] = rng.standard_normal(grid.shape) * s
    return out


def random_particles(grid, n, seed, flags=None, active_frac=1.0, pz_scale=0.5, q=-1e-15, margin=1.0):
    """Particles spread over the vacuum of the grid (avoiding metal when flags are given)."""
    rng = np.random.default_rng(seed)
    z = grid.zmin + grid.dz * (margin + rng.random(n) * (grid.nz - 1 - 2 * margin))
    r = grid.dr * (rng.random(n) * (grid.nr - 1 - margin))
    if flags is not None:
        for _ in range(50):
            iz = ((z - grid.zmin) / grid.dz).astype(np.int64)
            ir = (r / grid.dr).astype(np.int64)
            bad = (flags[iz, ir] & abi.FL_IS_METAL) | (flags[np.minimum(iz + 1, grid.nz), ir] & abi.FL_IS_METAL) | \
                  (flags[iz, np.minimum(ir + 1, grid.nr)] & abi.FL_IS_METAL)
            bad = bad != 0
            if not bad.any():
                break
            z[bad] = grid.zmin + grid.dz * (margin + rng.random(bad.sum()) * (grid.nz - 1 - 2 * margin))
            r[bad] = grid.dr * (rng.random(bad.sum()) * (grid.nr - 1 - margin))
    pz = rng.standard_normal(n) * pz_scale
    pr = rng.standard_normal(n) * pz_scale * 0.1
    rp = r * rng.standard_normal(n) * pz_scale * 0.05
    g = np.sqrt(pz * pz + pr * pr + np.where(r > 0, (rp / np.where(r > 0, r, 1)) ** 2, 0) + 1)
    st = np.where(rng.random(n) < active_frac, abi.PART_ACTIVE, 0).astype(np.uint32)
    return {
        "Q": np.full(n, q), "z": z, "r": r, "pz": pz, "pr": pr, "r_pphi": rp,
        "vz": K.c_mks * pz / g, "vr": K.c_mks * pr / g, "vphi": np.where(r > 0, K.c_mks * rp / np.where(r > 0, r, 1) / g, 0),
        "r0": r.copy(), "z0": z.copy(), "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(), "zmax": z.copy(),
        "zmin": z.copy(), "particleID": np.arange(n, dtype=np.int32), "generation": np.zeros(n, dtype=np.int16),
        "status": st,
    }


def bits(a):
    """Raw 64-bit patterns, so equality is bit-for-bit (it tells -0.0 from +0.0)."""
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)
