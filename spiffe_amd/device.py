"""Host-side mirror of the reference's per-step interface, over the C ABI.

The reference hands FIELDS/BEAM structs to plain C functions
(src/spiffe.h:474-561); here a `DeviceRun` owns the same state in HBM and
exposes the same operations by name: `advance_field_solutions`,
`zero_current_arrays`, `multiply_current_arrays`, `add_particle_currents`,
`advance_particles`, `add_poisson_correction`, `solve_poisson_cyl`,
`compute_imposed_Efield`, plus field/particle transfer in the reference's
[iz][ir] layout.  Every call goes to libspiffe_amd; nothing here computes
physics on the CPU.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import abi
from . import constants as K


class SpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("spiffe_amd error %d: %s" % (code, msg))
        self.code = code


@dataclass
class Grid:
    """define_geometry + integrate parameters (src/geometry.nl, src/integrate.nl)."""
    nz: int
    nr: int
    zmin: float
    zmax: float
    rmax: float
    dt: float
    modes: int = abi.TM_FIELDS | abi.RADIAL_INTERP | abi.LONGIT_INTERP
    lower_bc: int = abi.DIRICHLET
    upper_bc: int = abi.NEUMANN
    left_bc: int = abi.NEUMANN
    right_bc: int = abi.NEUMANN
    start_time: float = 0.0
    turn_off: tuple = (0, 0, 0, 0, 0, 0)
    constant_fields: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)

    @property
    def dz(self):
        return (self.zmax - self.zmin) / (self.nz - 1)   # src/geometry.c:75

    @property
    def dr(self):
        return self.rmax / (self.nr - 1)                 # src/geometry.c:76

    @property
    def shape(self):
        return (self.nz + 1, self.nr + 1)


def courant_dt(dz, dr, fraction=1.0):
    """fraction x the stability bound dtC = min(dz, dr)/c/sqrt(2), src/integrate.c:94."""
    return fraction * (min(dz, dr) / K.c_mks / np.sqrt(2.0))


class Context:
    """One HIP stream on one device; holds any number of runs."""

    def __init__(self, device=0):
        self.lib = abi.load()
        self.device = device
        h = ctypes.c_void_p()
        rc = self.lib.spf_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise SpfError(rc, "spf_ctx_create failed (no usable GPU?)")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise SpfError(rc, self.lib.spf_last_error(self.h).decode())

    def synchronize(self):
        self.check(self.lib.spf_synchronize(self.h))

    def stream(self):
        return self.lib.spf_stream(self.h)

    def timing(self, every=1):
        """Starts (and clears) the per-phase HIP-event timers, timing one call in `every`
        of each phase; every=0 stops them."""
        self.check(self.lib.spf_timing_enable(self.h, int(every)))

    def timings(self):
        """{phase: (ms over the timed launches, launches timed, all launches)}; synchronizing."""
        ms = (ctypes.c_double * len(abi.TIMERS))()
        cnt = (ctypes.c_int64 * len(abi.TIMERS))()
        tot = (ctypes.c_int64 * len(abi.TIMERS))()
        self.check(self.lib.spf_timing_read(self.h, ms, cnt, tot))
        return {name: (ms[i], cnt[i], tot[i]) for i, name in enumerate(abi.TIMERS)}

    def set_sor_shape(self, shape):
        """SOR tile shape of this context's solves: 0, 1, or -1 (by the points solved)."""
        self.check(self.lib.spf_set_sor_shape(self.h, int(shape)))

    def step_batch(self, runs):
        """One whole step of each run (DeviceRun objects of this context), their
        phases launched together over all of them."""
        ids = np.ascontiguousarray([r.id for r in runs], dtype=np.int32)
        self.check(self.lib.spf_step_batch(self.h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    def close(self):
        if self.h:
            self.lib.spf_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


SCREEN_ROW_DTYPE = np.dtype(abi.ScreenRow)
LOST_ROW_DTYPE = np.dtype(abi.LostRow)


class DeviceRun:
    def __init__(self, ctx, grid, flags, max_particles=0, screens=(), compat=True, QoMC=K.QoMC_electron,
                 stiffness=1.0, has_solenoid=False, screen_row_capacity=0, lost_row_capacity=0, cell_sort=0):
        """cell_sort=0 keeps the reference's particle order and sums (bit for bit);
        cell_sort=N re-sorts the particles by cell every N steps and deposits in a
        fixed order of the library's own (needs compat=False)."""
        self.ctx, self.lib, self.grid = ctx, ctx.lib, grid
        flags = np.ascontiguousarray(flags, dtype=np.int16)
        if flags.shape != grid.shape:
            raise ValueError("flags must have shape (nz+1, nr+1)")
        sz = np.ascontiguousarray([s[0] for s in screens], dtype=np.float64)
        sd = np.ascontiguousarray([s[1] for s in screens], dtype=np.int32)
        d = abi.RunDesc()
        d.nz, d.nr = grid.nz, grid.nr
        d.zmin, d.zmax, d.rmax = grid.zmin, grid.zmax, grid.rmax
        d.dz, d.dr = grid.dz, grid.dr
        d.lower_bc, d.upper_bc, d.left_bc, d.right_bc = grid.lower_bc, grid.upper_bc, grid.left_bc, grid.right_bc
        d.modes = grid.modes
        d.dt, d.start_time = grid.dt, grid.start_time
        d.epsilon_o, d.mu_o, d.c_mks = K.epsilon_o, K.mu_o, K.c_mks
        d.QoMC, d.stiffness = QoMC, stiffness
        for k in range(6):
            d.turn_off[k] = int(grid.turn_off[k])
            d.constant_fields[k] = float(grid.constant_fields[k])
        d.has_solenoid = 1 if has_solenoid else 0
        d.metal_flags = _ptr(flags, ctypes.c_int16)
        d.max_particles = int(max_particles)
        d.n_screens = len(screens)
        d.screen_z = _ptr(sz, ctypes.c_double)
        d.screen_direction = _ptr(sd, ctypes.c_int32)
        d.compat = 1 if compat else 0
        d.screen_row_capacity = screen_row_capacity
        d.lost_row_capacity = lost_row_capacity
        d.cell_sort = int(cell_sort)
        self.desc = d
        rid = ctypes.c_int32()
        ctx.check(self.lib.spf_run_create(ctx.h, ctypes.byref(d), ctypes.byref(rid)))
        self.id = rid.value
        self.flags = flags

    # ---- lifecycle / transfer ----
    def close(self):
        if self.id is not None and self.ctx.h:
            self.ctx.check(self.lib.spf_run_destroy(self.ctx.h, self.id))
        self.id = None

    def footprint(self):
        return self.lib.spf_run_footprint(ctypes.byref(self.desc))

    def move_to(self, ctx):
        """Hands the run to another context of the same device (spf_run_move)."""
        rid = ctypes.c_int32()
        self.ctx.check(self.lib.spf_run_move(self.ctx.h, self.id, ctx.h, ctypes.byref(rid)))
        self.ctx, self.id = ctx, rid.value

    @staticmethod
    def footprint_for(grid, max_particles, n_screens, cell_sort=0):
        """HBM bytes a run of this grid, capacity and screen count would hold
        (spf_run_footprint), before creating it."""
        d = abi.RunDesc()
        d.nz, d.nr = grid.nz, grid.nr
        d.max_particles = int(max_particles)
        d.n_screens = int(n_screens)
        d.cell_sort = int(cell_sort)
        return abi.load().spf_run_footprint(ctypes.byref(d))

    def sort_particles(self):
        """Re-sorts a cell-sort run's particles by cell now (spf_particles_sort)."""
        self.ctx.check(self.lib.spf_particles_sort(self.ctx.h, self.id))

    def set_cell_sort(self, interval):
        """Changes a cell-sort run's re-sort interval (spf_set_cell_sort)."""
        self.ctx.check(self.lib.spf_set_cell_sort(self.ctx.h, self.id, int(interval)))

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.grid.shape:
            raise ValueError("field must have shape (nz+1, nr+1)")
        self.ctx.check(self.lib.spf_field_upload(self.ctx.h, self.id, abi.FIELD_ID[name], _ptr(a, ctypes.c_double)))

    def download(self, name):
        a = np.empty(self.grid.shape, dtype=np.float64)
        self.ctx.check(self.lib.spf_field_download(self.ctx.h, self.id, abi.FIELD_ID[name], _ptr(a, ctypes.c_double)))
        return a

    def set_step(self, step):
        self.ctx.check(self.lib.spf_set_step(self.ctx.h, self.id, int(step)))

    def set_z_forces_start(self, z):
        """integrate's z_forces_start, used by step() (src/integrate.c:186-188)."""
        self.ctx.check(self.lib.spf_set_z_forces_start(self.ctx.h, self.id, float(z)))

    def set_j_filter(self, multiplier):
        """integrate's J_filter_multiplier, used by step() (src/integrate.c:110-116, :168-178)."""
        self.ctx.check(self.lib.spf_set_j_filter(self.ctx.h, self.id, float(multiplier)))

    def get_step(self):
        s, t = ctypes.c_int64(), ctypes.c_double()
        self.ctx.check(self.lib.spf_get_step(self.ctx.h, self.id, ctypes.byref(s), ctypes.byref(t)))
        return s.value, t.value

    def append(self, parts):
        n = len(parts["z"])
        cols = {}
        p = abi.Particles()
        for name in abi.PARTICLE_DOUBLES:
            cols[name] = np.ascontiguousarray(parts[name], dtype=np.float64)
            setattr(p, name, _ptr(cols[name], ctypes.c_double))
        cols["particleID"] = np.ascontiguousarray(parts["particleID"], dtype=np.int32)
        cols["generation"] = np.ascontiguousarray(parts["generation"], dtype=np.int16)
        cols["status"] = np.ascontiguousarray(parts["status"], dtype=np.uint32)
        p.particleID = _ptr(cols["particleID"], ctypes.c_int32)
        p.generation = _ptr(cols["generation"], ctypes.c_int16)
        p.status = _ptr(cols["status"], ctypes.c_uint32)
        self.ctx.check(self.lib.spf_particles_append(self.ctx.h, self.id, n, ctypes.byref(p)))
        self.ctx.synchronize()

    def count(self):
        n, a = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.lib.spf_particles_count(self.ctx.h, self.id, ctypes.byref(n), ctypes.byref(a)))
        return n.value, a.value

    def particles(self):
        n, _ = self.count()
        out = {name: np.empty(n, dtype=np.float64) for name in abi.PARTICLE_DOUBLES}
        out["particleID"] = np.empty(n, dtype=np.int32)
        out["generation"] = np.empty(n, dtype=np.int16)
        out["status"] = np.empty(n, dtype=np.uint32)
        p = abi.Particles()
        for name in abi.PARTICLE_DOUBLES:
            setattr(p, name, _ptr(out[name], ctypes.c_double))
        p.particleID = _ptr(out["particleID"], ctypes.c_int32)
        p.generation = _ptr(out["generation"], ctypes.c_int16)
        p.status = _ptr(out["status"], ctypes.c_uint32)
        got = ctypes.c_int64()
        self.ctx.check(self.lib.spf_particles_download(self.ctx.h, self.id, ctypes.byref(p), n, ctypes.byref(got)))
        return out

    def drain(self, max_rows=1 << 20, max_lost=1 << 20):
        rows = np.zeros(max_rows, dtype=SCREEN_ROW_DTYPE)
        lost = np.zeros(max_lost, dtype=LOST_ROW_DTYPE)
        nr_, nl_ = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.lib.spf_events_drain(self.ctx.h, self.id, _ptr(rows, abi.ScreenRow), max_rows,
                                                 ctypes.byref(nr_), _ptr(lost, abi.LostRow), max_lost,
                                                 ctypes.byref(nl_)))
        return rows[:nr_.value], lost[:nl_.value]

    # ---- imposed solenoid field (src/solenoid.c) ----
    def define_solenoid(self, **namelist):
        """process_solenoid_definition with define_solenoid values (src/solenoid.nl),
        plus the gaussian / gaussian_center / gaussian_sigma extension.  Returns the
        ring current applied."""
        d = abi.SolenoidDesc()
        self.lib.spf_solenoid_defaults(ctypes.byref(d))
        names = {n for n, _ in abi.SolenoidDesc._fields_}
        for key, val in namelist.items():
            if key not in names:
                raise ValueError("unknown define_solenoid parameter %r" % key)
            setattr(d, key, val)
        cur = ctypes.c_double()
        self.ctx.check(self.lib.spf_solenoid_add(self.ctx.h, self.id, ctypes.byref(d), ctypes.byref(cur)))
        return cur.value

    # ---- cathode emission (src/cathode.c) ----
    def define_cathode(self, profile=None, **namelist):
        """process_cathode_definition: `namelist` holds define_cathode values
        (src/cathode.nl); `profile` is (time, factor) arrays.  Returns the
        indices of the cathodes added."""
        d = abi.CathodeDesc()
        self.lib.spf_cathode_defaults(ctypes.byref(d))
        d.e_mks, d.me_mks, d.k_boltzmann_mks = K.e_mks, K.me_mks, K.k_boltzmann_mks
        names = {n for n, _ in abi.CATHODE_FIELDS}
        for key, val in namelist.items():
            if key not in names or key in ("n_profile", "profile_time", "profile_factor"):
                raise ValueError("unknown define_cathode parameter %r" % key)
            setattr(d, key, val)
        keep = []
        if profile is not None:
            t = np.ascontiguousarray(profile[0], dtype=np.float64)
            a = np.ascontiguousarray(profile[1], dtype=np.float64)
            if t.shape != a.shape or t.ndim != 1:
                raise ValueError("profile time and factor must be 1-D arrays of one length")
            keep = [t, a]
            d.n_profile = len(t)
            d.profile_time, d.profile_factor = _ptr(t, ctypes.c_double), _ptr(a, ctypes.c_double)
        first, count = ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(self.lib.spf_cathode_define(self.ctx.h, self.id, ctypes.byref(d), ctypes.byref(first),
                                                   ctypes.byref(count)))
        del keep
        return list(range(first.value, first.value + count.value))

    def emit_electrons_from_cathode(self, cathode):
        n = ctypes.c_int64()
        self.ctx.check(self.lib.spf_cathode_emit(self.ctx.h, self.id, int(cathode), ctypes.byref(n)))
        return n.value

    def set_current_density(self, cathode):
        self.ctx.check(self.lib.spf_cathode_set_current_density(self.ctx.h, self.id, int(cathode)))

    def cathode_state(self, cathode):
        st = abi.CathodeState()
        self.ctx.check(self.lib.spf_cathode_state_get(self.ctx.h, self.id, int(cathode), ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in abi.CathodeState._fields_}

    def next_particle_id(self, set_next=0):
        v = ctypes.c_int64()
        self.ctx.check(self.lib.spf_particle_id_next(self.ctx.h, self.id, int(set_next), ctypes.byref(v)))
        return v.value

    # ---- the reference's per-step operations ----
    def advance_field_solutions(self):
        self.ctx.check(self.lib.spf_fields_advance(self.ctx.h, self.id))

    def zero_current_arrays(self):
        self.ctx.check(self.lib.spf_currents_reset(self.ctx.h, self.id))

    def multiply_current_arrays(self, factor):
        self.ctx.check(self.lib.spf_currents_scale(self.ctx.h, self.id, float(factor)))

    def add_line_current(self, direction, iposition, istart, iend, J):
        self.ctx.check(self.lib.spf_currents_add_line(self.ctx.h, self.id, int(direction), int(iposition),
                                                      int(istart), int(iend), float(J)))

    def add_particle_currents(self):
        self.ctx.check(self.lib.spf_currents_deposit(self.ctx.h, self.id))

    def advance_particles(self, z_forces_start=-1.0):
        self.ctx.check(self.lib.spf_particles_push(self.ctx.h, self.id, float(z_forces_start)))

    def setup_poisson_correction(self, step_interval, accuracy=1e-6, maximum_iterations=1000, start_time=0.0,
                                 error_charge_threshold=0.0, guess_code=0):
        p = abi.PoissonParams(start_time, step_interval, accuracy, error_charge_threshold, maximum_iterations,
                              guess_code)
        self.ctx.check(self.lib.spf_poisson_setup(self.ctx.h, self.id, ctypes.byref(p)))

    def add_poisson_correction(self, space_charge):
        it = ctypes.c_int64()
        self.ctx.check(self.lib.spf_poisson_correct(self.ctx.h, self.id, 1 if space_charge else 0, ctypes.byref(it)))
        return it.value

    def sor_batch_half_sweeps(self):
        """Half-sweeps one SOR batch launch of this run performs (its tile shape)."""
        v = ctypes.c_int32()
        self.ctx.check(self.lib.spf_sor_batch_half_sweeps(self.ctx.h, self.id, ctypes.byref(v)))
        return v.value

    def poisson_stats(self):
        """(result of the last correction that solved, corrections solved, half-sweeps in all)."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.lib.spf_poisson_stats(self.ctx.h, self.id, ctypes.byref(a), ctypes.byref(b),
                                                  ctypes.byref(c)))
        return a.value, b.value, c.value

    def solve_poisson_cyl(self, which, accuracy, maximum_iterations):
        it = ctypes.c_int64()
        self.ctx.check(self.lib.spf_poisson_solve(self.ctx.h, self.id, int(which), float(accuracy),
                                                  int(maximum_iterations), ctypes.byref(it)))
        return it.value

    def poisson_sweeps(self, which, n0, count):
        """`count` half-sweeps with no stopping test (stream-ordered)."""
        self.ctx.check(self.lib.spf_poisson_sweeps(self.ctx.h, self.id, int(which), int(n0), int(count)))

    def compute_imposed_Efield(self, initialize, ramp_time=0.0, flat_top_time=0.0):
        self.ctx.check(self.lib.spf_imposed_efield(self.ctx.h, self.id, 1 if initialize else 0, float(ramp_time),
                                                   float(flat_top_time)))

    def step(self, lines=()):
        """One loop body of src/integrate.c:134-194, emission from the run's cathodes included."""
        n = len(lines)
        d = np.ascontiguousarray([l[0] for l in lines], dtype=np.int32)
        pos = np.ascontiguousarray([l[1] for l in lines], dtype=np.int64)
        st = np.ascontiguousarray([l[2] for l in lines], dtype=np.int64)
        en = np.ascontiguousarray([l[3] for l in lines], dtype=np.int64)
        J = np.ascontiguousarray([l[4] for l in lines], dtype=np.float64)
        self.ctx.check(self.lib.spf_step(self.ctx.h, self.id, n, _ptr(d, ctypes.c_int32), _ptr(pos, ctypes.c_int64),
                                         _ptr(st, ctypes.c_int64), _ptr(en, ctypes.c_int64),
                                         _ptr(J, ctypes.c_double)))
