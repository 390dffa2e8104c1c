"""Synthetic workloads of the benchmark configurations (seeded).

gun_hires(): the high-resolution single run -- the same 1.6-cell gun at
dr = dz = 25 um (4201 x 1681 grid lines, 7.06M cells), 16M macroparticles
carrying 100 pC in a Gaussian bunch (sigma_t 3 ps) of radius 0.6 mm, TM+TE,
the Gaussian solenoid, Poisson correction every 5 steps to 1e-8.

gun_16cell(): the production-size single run -- a 1.6-cell S-band gun
(half cell 0-3.5 cm, full cell 3.5-10.5 cm, cell radius 4.2 cm, iris radius
1.25 cm, iris 0.5 cm thick), dr = dz = 0.1 mm (1051 x 421 grid lines),
TM+TE fields, an external Gaussian solenoid (0.25 T at z = 20 cm, sigma
5 cm), 1M macroparticles carrying 1 nC, Poisson correction every 10 steps to
1e-8, a screen at z = 10 cm.  Fields start from a synthetic
standing-wave pattern scaled to 120 MV/m on axis and the bunch starts as a
uniform 1 mm-radius slug just off the cathode: the shapes and sizes of the
configuration with synthetic contents (no eigenmode solve, no emission
history).
"""
import os

import numpy as np

from . import abi
from . import constants as K
from .device import Grid, courant_dt
from .geometry import boundary_file

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
GUN_GEO = os.path.join(DECKS, "gun_1p6cell.geo")
DC_GUN_GEO = os.path.join(DECKS, "dc_gun.geo")
GUN_POLY_16 = [(0.0, 0.0), (0.0, 0.042), (0.0325, 0.042), (0.0325, 0.0125), (0.0375, 0.0125), (0.0375, 0.042),
               (0.105, 0.042), (0.105, 0.0), (0.0, 0.0)]   # the outline of gun_1p6cell.geo


def gun_grid(dz=1e-4, modes=abi.TM_FIELDS | abi.TE_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP |
             abi.SPACE_CHARGE):
    zmax, rmax = 0.105, 0.042
    nz = int(round(zmax / dz)) + 1
    nr = int(round(rmax / dz)) + 1
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=zmax, rmax=rmax, dt=0.0, modes=modes)
    g.dt = courant_dt(g.dz, g.dr, 0.9)
    return g


def gun_flags(grid):
    """Flags of the 1.6-cell gun boundary file drawn on `grid`."""
    return boundary_file(GUN_GEO, grid.nz, grid.nr, grid.zmin, grid.zmax, grid.rmax)[0]


def dc_gun_grid(dz=5e-5, modes=abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP | abi.SPACE_CHARGE):
    """The DC-gun request of the serving stream: 0.05 mm mesh over 3 cm x 1.2 cm."""
    zmax, rmax = 0.030, 0.012
    nz = int(round(zmax / dz)) + 1
    nr = int(round(rmax / dz)) + 1
    g = Grid(nz=nz, nr=nr, zmin=0.0, zmax=zmax, rmax=rmax, dt=0.0, modes=modes)
    g.dt = courant_dt(g.dz, g.dr, 0.9)
    return g


def dc_gun_geometry(grid):
    """(flags, PsiImposed boundary potentials) of the DC-gun boundary file on `grid`."""
    flags, psi, _ = boundary_file(DC_GUN_GEO, grid.nz, grid.nr, grid.zmin, grid.zmax, grid.rmax)
    return flags, psi


def gun_fields(grid, E0=120e6):
    """A standing-wave Ez pattern, pi-mode-like between the two cells, in the
    phase that accelerates electrons off the cathode (Ez < 0 in the half cell)."""
    z = grid.zmin + grid.dz * (np.arange(grid.nz + 1) - 0.5)
    r = grid.dr * np.arange(grid.nr + 1)
    from scipy.special import j0
    ez_z = np.where(z < 0.035, -1.0, 1.0) * np.abs(np.cos(np.pi * (z - 0.035) / 0.07))
    Ez = E0 * np.outer(ez_z, j0(2.404825557695773 * np.minimum(r, 0.042) / 0.042))
    return {"Ez": Ez}


def gun_bunch(grid, n, charge=1e-9, radius=1e-3, seed=7, z0=0.3e-3, length=1.5e-3, pz=0.6):
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.random(n))
    z = grid.zmin + z0 + length * rng.random(n)
    pzv = pz * (1 + 0.02 * rng.standard_normal(n))
    prv = 0.01 * rng.standard_normal(n)
    rp = np.zeros(n)
    g = np.sqrt(pzv * pzv + prv * prv + 1)
    q = -charge / n
    return {"Q": np.full(n, q), "z": z, "r": r, "pz": pzv, "pr": prv, "r_pphi": rp,
            "vz": K.c_mks * pzv / g, "vr": K.c_mks * prv / g, "vphi": np.zeros(n),
            "r0": r.copy(), "z0": z.copy(), "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(),
            "zmax": z.copy(), "zmin": z.copy(), "particleID": np.arange(n, dtype=np.int32),
            "generation": np.zeros(n, dtype=np.int16), "status": np.full(n, abi.PART_ACTIVE, dtype=np.uint32)}


def gun_bunch_gaussian(grid, n, charge=100e-12, radius=0.6e-3, sigma_t=3e-12, seed=11, z_center=2.0e-3, pz=0.6):
    """A bunch with a Gaussian longitudinal profile of rms duration sigma_t
    (sigma_z = beta c sigma_t at momentum pz), uniform over a disc of the given
    radius, just off the cathode; the tails are cut at 3 sigma."""
    rng = np.random.default_rng(seed)
    beta = pz / np.sqrt(1 + pz * pz)
    sigma_z = beta * K.c_mks * sigma_t
    u = rng.standard_normal(n)
    u = np.where(np.abs(u) > 3, rng.uniform(-3, 3, n), u)
    z = grid.zmin + z_center + sigma_z * u
    r = radius * np.sqrt(rng.random(n))
    pzv = pz * (1 + 0.02 * rng.standard_normal(n))
    prv = 0.01 * rng.standard_normal(n)
    g = np.sqrt(pzv * pzv + prv * prv + 1)
    q = -charge / n
    return {"Q": np.full(n, q), "z": z, "r": r, "pz": pzv, "pr": prv, "r_pphi": np.zeros(n),
            "vz": K.c_mks * pzv / g, "vr": K.c_mks * prv / g, "vphi": np.zeros(n),
            "r0": r.copy(), "z0": z.copy(), "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(),
            "zmax": z.copy(), "zmin": z.copy(), "particleID": np.arange(n, dtype=np.int32),
            "generation": np.zeros(n, dtype=np.int16), "status": np.full(n, abi.PART_ACTIVE, dtype=np.uint32)}


def gun_leavers(grid, n, seed=6):
    """Particles that leave during the first steps: half drift back through the
    cathode plane z = zmin, half out through the wall at the top of the half
    cell (ids from 100 000 000 up)."""
    rng = np.random.default_rng(seed)
    h = n // 2
    z = np.concatenate([grid.zmin + rng.uniform(2e-6, 60e-6, h), grid.zmin + rng.uniform(5e-3, 30e-3, n - h)])
    r = np.concatenate([rng.uniform(0, 1e-3, h), rng.uniform(41.80e-3, 41.99e-3, n - h)])
    pz = np.concatenate([-0.5 * np.ones(h), 0.1 * rng.standard_normal(n - h)])
    pr = np.concatenate([0.01 * rng.standard_normal(h), 0.8 * np.ones(n - h)])
    gam = np.sqrt(pz * pz + pr * pr + 1)
    return {"Q": np.full(n, -1e-17), "z": z, "r": r, "pz": pz, "pr": pr, "r_pphi": np.zeros(n),
            "vz": K.c_mks * pz / gam, "vr": K.c_mks * pr / gam, "vphi": np.zeros(n), "r0": r.copy(),
            "z0": z.copy(), "t0": np.zeros(n), "rmax": r.copy(), "rmin": r.copy(), "zmax": z.copy(),
            "zmin": z.copy(), "particleID": np.arange(n, dtype=np.int32) + 100_000_000,
            "generation": np.zeros(n, dtype=np.int16), "status": np.full(n, abi.PART_ACTIVE, dtype=np.uint32)}


GUN_SCREENS = [(0.10, 1)]
# a screen inside the high-resolution bunch (centred 2 mm from the cathode):
# about 0.7% of its particles cross it every step
HIRES_EVENT_SCREENS = [(0.002, 1), (0.10, 1)]
# external solenoid: Gaussian on-axis Bz, 0.25 T peak at z = 20 cm, sigma 5 cm
GUN_SOLENOID = dict(gaussian=1, Bz_peak=0.25, gaussian_center=0.20, gaussian_sigma=0.05)
GUN_POISSON = dict(step_interval=10, accuracy=1e-8, maximum_iterations=3000)
# the high-resolution run: 25 um mesh, 16M particles, 100 pC, correction every
# 5 steps to 1e-8 (about 19 000 half-sweeps per correction once Psi is warm;
# the cap is there for the first, cold solve)
HIRES_DZ = 25e-6
HIRES_PARTICLES = 16_000_000
HIRES_CHARGE = 100e-12
HIRES_RADIUS = 0.6e-3
HIRES_POISSON = dict(step_interval=5, accuracy=1e-8, maximum_iterations=40000)
