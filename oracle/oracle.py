"""Python handle on the CPU restatement (oracle/libspf_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's CPU-baseline
leg use this as the checker.  The product path (spiffe_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libspf_oracle.so")
LIB_OMP = os.path.join(HERE, "libspf_oracle_omp.so")   # the same sources with OpenMP (CPU baseline only)

_d, _dp = ctypes.c_double, ctypes.POINTER(ctypes.c_double)
_i16p = ctypes.POINTER(ctypes.c_int16)


class OFields(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("nr", ctypes.c_int32),
                ("zmin", _d), ("zmax", _d), ("rmax", _d), ("dz", _d), ("dr", _d),
                ("lower_bc", ctypes.c_int32), ("upper_bc", ctypes.c_int32),
                ("right_bc", ctypes.c_int32), ("left_bc", ctypes.c_int32),
                ("start_time", _d), ("time", _d), ("dti", _d),
                ("time_step", ctypes.c_int64), ("modes", ctypes.c_uint32)] + \
               [(n, ctypes.c_int16) for n in ["turnOffEz", "turnOffEr", "turnOffEphi", "turnOffBz", "turnOffBr",
                                               "turnOffBphi"]] + \
               [(n, _d) for n in ["constantEz", "constantEr", "constantEphi", "constantBz", "constantBr",
                                  "constantBphi", "epsilon_o", "mu_o", "c_mks"]] + \
               [(n, _dp) for n in ["Ez", "Jz", "Er", "Jr", "Bphi", "Bphi2", "Ephi", "Jphi", "Br", "Br2", "Bz",
                                   "Bz2"]] + \
               [("metal_flags", _i16p)] + \
               [(n, _dp) for n in ["EzImposed", "ErImposed", "PsiImposed", "BzImposed", "BrImposed"]] + \
               [("imposedEFieldRampTime", _d), ("imposedEFieldFlatTopTime", _d), ("Psi", _dp), ("Q", _dp)]


PDOUBLES = ["Q", "z", "r", "pz", "pr", "r_pphi", "vz", "vr", "vphi", "r0", "z0", "t0", "rmax", "rmin", "zmax",
            "zmin"]


class OBeam(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64), ("npActive", ctypes.c_int64), ("max_np", ctypes.c_int64),
                ("QoMC", _d), ("stiffness", _d)] + [(n, _dp) for n in PDOUBLES] + \
               [("particleID", ctypes.POINTER(ctypes.c_int32)), ("generation", _i16p),
                ("status", ctypes.POINTER(ctypes.c_uint32))]


class OScreen(ctypes.Structure):
    _fields_ = [("z_position", _d), ("direction", ctypes.c_int32), ("pad", ctypes.c_int32)]


class OScreenRow(ctypes.Structure):
    _fields_ = [(n, _d) for n in ["r", "pz", "pr", "pphi", "t", "q", "r0", "t0", "rmin", "rmax", "zmin",
                                  "zmax"]] + \
               [("ip", ctypes.c_int64), ("particleID", ctypes.c_int32), ("generation", ctypes.c_int16),
                ("screen", ctypes.c_int16)]


class OLostRow(ctypes.Structure):
    _fields_ = [("t", _d)] + [(n, ctypes.c_float) for n in ["rLost", "zLost", "pz", "pr", "r0", "z0", "q"]] + \
               [("particleID", ctypes.c_int32), ("ip", ctypes.c_int64), ("generation", ctypes.c_int16),
                ("directionCode", ctypes.c_int16), ("pad", ctypes.c_int32)]


class OPoisson(ctypes.Structure):
    _fields_ = [("start_time", _d), ("step_interval", ctypes.c_int64), ("accuracy", _d),
                ("error_charge_threshold", _d), ("maximum_iterations", ctypes.c_int64),
                ("guess_code", ctypes.c_int64), ("pc_start_step", ctypes.c_int64),
                ("last_iterations", ctypes.c_int64), ("Qreal", _d), ("Qerror", _d), ("Qsurface", _d)]


class OCathodeDesc(ctypes.Structure):
    """ocathode_desc: the define_cathode namelist (src/cathode.nl)."""
    _fields_ = [
        ("z_position", _d), ("inner_radius", _d), ("outer_radius", _d),
        ("current_density", _d), ("temperature", _d), ("work_function", _d),
        ("field_emission", ctypes.c_int32), ("add_thermal_velocities", ctypes.c_int32),
        ("field_emission_beta", _d),
        ("determine_temperature", ctypes.c_int32), ("autophase", ctypes.c_int32),
        ("electrons_per_macroparticle", _d), ("number_per_step", _d),
        ("start_time", _d), ("stop_time", _d), ("time_offset", _d),
        ("initial_pz", _d), ("initial_omega", _d), ("stiffness", _d),
        ("discretize_radii", ctypes.c_int32), ("spread_over_dt", ctypes.c_int32),
        ("zoned_emission", ctypes.c_int32), ("distribution_correction_interval", ctypes.c_int32),
        ("halton_radix_dt", ctypes.c_int32), ("halton_radix_r", ctypes.c_int32),
        ("random_number_seed", ctypes.c_int64),
        ("n_profile", ctypes.c_int64), ("profile_time", _dp), ("profile_factor", _dp),
        ("e_mks", _d), ("me_mks", _d), ("k_boltzmann_mks", _d)]


class OSolenoidDesc(ctypes.Structure):
    _fields_ = [(n, _d) for n in ["radius", "evaluation_radius_limit", "z_start", "z_end", "current", "Bz_peak"]] + \
               [("turns", ctypes.c_int64), ("symmetry", ctypes.c_int64), ("bucking", ctypes.c_int32),
                ("z_buck", _d), ("gaussian", ctypes.c_int32), ("gaussian_center", _d), ("gaussian_sigma", _d)]


class ORng(ctypes.Structure):
    _fields_ = [("key0", ctypes.c_uint64), ("block", ctypes.c_uint64), ("buf", ctypes.c_uint64 * 4),
                ("pos", ctypes.c_int32), ("has_saved", ctypes.c_int32), ("saved", _d)]


SCREEN_ROW_DTYPE = np.dtype(OScreenRow)
LOST_ROW_DTYPE = np.dtype(OLostRow)
FIELD_NAMES = ["Ez", "Jz", "Er", "Jr", "Bphi", "Bphi2", "Ephi", "Jphi", "Br", "Br2", "Bz", "Bz2",
               "EzImposed", "ErImposed", "PsiImposed", "BzImposed", "BrImposed", "Psi", "Q"]

_lib = None
_lib_omp = None


def build():
    """Compiles the restatement with the committed Makefile (gcc, -ffp-contract=off)."""
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def load(threaded=False):
    """The serial restatement (the checker), or with threaded=True its OpenMP
    build (timing of the all-cores CPU baseline)."""
    global _lib, _lib_omp
    if threaded:
        if _lib_omp is None:
            if not os.path.exists(LIB_OMP):
                build()
            _lib_omp = _prototypes(ctypes.CDLL(LIB_OMP))
        return _lib_omp
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _prototypes(ctypes.CDLL(LIB))
    return _lib


def _prototypes(lib):
    P = ctypes.POINTER
    lib.oracle_calculate_flags.argtypes = [_i16p, ctypes.c_int32, ctypes.c_int32]
    lib.oracle_advance_field_solutions.argtypes = [P(OFields)]
    lib.oracle_zero_current_arrays.argtypes = [P(OFields)]
    lib.oracle_multiply_current_arrays.argtypes = [P(OFields), _d]
    lib.oracle_add_line_current.argtypes = [P(OFields), ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, _d]
    lib.oracle_add_particle_currents.argtypes = [P(OFields), P(OBeam)]
    lib.oracle_check_loss_code.argtypes = [ctypes.c_int64, ctypes.c_int64, P(OFields)]
    lib.oracle_check_loss_code.restype = ctypes.c_short
    lib.oracle_interpolate_fields.argtypes = [_dp] * 6 + [_d, _d, P(OFields)]
    lib.oracle_find_loss_coordinates.argtypes = [_dp, _dp, P(ctypes.c_short), _d, _d, _d, _d, P(OFields)]
    lib.oracle_advance_particles.argtypes = [P(OFields), P(OBeam), P(OScreen), ctypes.c_int32, _d,
                                             ctypes.c_int32, P(OScreenRow), ctypes.c_int64,
                                             P(ctypes.c_int64), P(OLostRow), ctypes.c_int64,
                                             P(ctypes.c_int64)]
    lib.oracle_solve_poisson_cyl.argtypes = [_dp, _dp, _i16p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                             ctypes.c_int64, _d, _d, _d, ctypes.c_int64, _dp]
    lib.oracle_solve_poisson_cyl.restype = ctypes.c_int64
    lib.oracle_poisson_omega_schedule.argtypes = [ctypes.c_int64, ctypes.c_int64, _d, _d, ctypes.c_int64, _dp]
    lib.oracle_add_poisson_correction.argtypes = [P(OFields), P(OBeam), ctypes.c_int32, P(OPoisson)]
    lib.oracle_compute_imposed_efield.argtypes = [P(OFields), ctypes.c_int32, P(OPoisson), _dp]
    lib.oracle_antenna_current.argtypes = [_dp, _dp, ctypes.c_int64, P(ctypes.c_int64), _d, _d, _d, _d,
                                           P(ctypes.c_int32)]
    lib.oracle_antenna_current.restype = _d
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.oracle_philox4x64_10.argtypes = [u64p, u64p, u64p]
    lib.oracle_rng_seed.argtypes = [P(ORng), ctypes.c_uint64]
    lib.oracle_rng_next64.argtypes = [P(ORng)]
    lib.oracle_rng_next64.restype = ctypes.c_uint64
    lib.oracle_rng_uniform.argtypes = [P(ORng)]
    lib.oracle_rng_uniform.restype = _d
    lib.oracle_rng_gauss.argtypes = [P(ORng)]
    lib.oracle_rng_gauss.restype = _d
    lib.oracle_radical_inverse.argtypes = [ctypes.c_uint64, ctypes.c_int32]
    lib.oracle_radical_inverse.restype = _d
    lib.oracle_cathodes_create.restype = ctypes.c_void_p
    lib.oracle_cathodes_destroy.argtypes = [ctypes.c_void_p]
    lib.oracle_cathodes_rng.argtypes = [ctypes.c_void_p]
    lib.oracle_cathodes_rng.restype = P(ORng)
    lib.oracle_cathodes_next_id.argtypes = [ctypes.c_void_p]
    lib.oracle_cathodes_next_id.restype = ctypes.c_int64
    lib.oracle_cathode_define.argtypes = [ctypes.c_void_p, P(OFields), P(OCathodeDesc), P(ctypes.c_int32),
                                          P(ctypes.c_int32), ctypes.c_char_p, ctypes.c_int64]
    lib.oracle_field_emission_current_density.argtypes = [_d, _d, _d]
    lib.oracle_field_emission_current_density.restype = _d
    lib.oracle_set_current_density.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(OFields)]
    lib.oracle_emit_electrons_from_cathode.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(OBeam), P(OFields),
                                                       ctypes.c_char_p, ctypes.c_int64]
    lib.oracle_cathode_state.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, P(ctypes.c_int64)]
    lib.oracle_ellip_ke.argtypes = [_d, _dp, _dp]
    lib.oracle_b_ring.argtypes = [_dp, _dp, _d, _d, _d]
    lib.oracle_b_solenoid.argtypes = [_dp, _dp, _d, _d, _d, _d, _d, ctypes.c_int64, ctypes.c_int64]
    lib.oracle_define_solenoid.argtypes = [P(OFields), P(OSolenoidDesc), _dp, ctypes.c_char_p, ctypes.c_int64]
    return lib


def derive_flags(marks):
    """The C restatement of src/geometry.c:319-362 applied to a copy of `marks`."""
    f = np.ascontiguousarray(marks, dtype=np.int16).copy()
    load().oracle_calculate_flags(f.ctypes.data_as(_i16p), f.shape[0] - 1, f.shape[1] - 1)
    return f


class OracleSim:
    """FIELDS + BEAM state of the restatement in numpy arrays ([iz][ir] layout)."""

    def __init__(self, grid, flags, max_particles=0, QoMC=None, stiffness=1.0, has_solenoid=False, constants=None,
                 threaded=False):
        from spiffe_amd import constants as K   # the same constants the product passes to the device
        self.lib = load(threaded)
        self.grid = grid
        shape = grid.shape
        self.arr = {n: np.zeros(shape, dtype=np.float64) for n in FIELD_NAMES}
        self.flags = np.ascontiguousarray(flags, dtype=np.int16).copy()
        f = OFields()
        f.nz, f.nr = grid.nz, grid.nr
        f.zmin, f.zmax, f.rmax, f.dz, f.dr = grid.zmin, grid.zmax, grid.rmax, grid.dz, grid.dr
        f.lower_bc, f.upper_bc, f.right_bc, f.left_bc = grid.lower_bc, grid.upper_bc, grid.right_bc, grid.left_bc
        f.start_time, f.time, f.dti, f.time_step, f.modes = grid.start_time, 0.0, grid.dt, 0, grid.modes
        for k, n in enumerate(["turnOffEz", "turnOffEr", "turnOffEphi", "turnOffBz", "turnOffBr", "turnOffBphi"]):
            setattr(f, n, int(grid.turn_off[k]))
        for k, n in enumerate(["constantEz", "constantEr", "constantEphi", "constantBz", "constantBr",
                               "constantBphi"]):
            setattr(f, n, float(grid.constant_fields[k]))
        f.epsilon_o, f.mu_o, f.c_mks = K.epsilon_o, K.mu_o, K.c_mks
        for n in FIELD_NAMES:
            if n in ("BzImposed", "BrImposed") and not has_solenoid:
                setattr(f, n, ctypes.cast(None, _dp))
            else:
                setattr(f, n, self.arr[n].ctypes.data_as(_dp))
        f.metal_flags = self.flags.ctypes.data_as(_i16p)
        self.f = f
        cap = max(int(max_particles), 1)
        self.p = {n: np.zeros(cap, dtype=np.float64) for n in PDOUBLES}
        self.p["particleID"] = np.zeros(cap, dtype=np.int32)
        self.p["generation"] = np.zeros(cap, dtype=np.int16)
        self.p["status"] = np.zeros(cap, dtype=np.uint32)
        b = OBeam()
        b.np, b.npActive, b.max_np = 0, 0, cap
        b.QoMC = K.QoMC_electron if QoMC is None else QoMC
        b.stiffness = stiffness
        for n in PDOUBLES:
            setattr(b, n, self.p[n].ctypes.data_as(_dp))
        b.particleID = self.p["particleID"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        b.generation = self.p["generation"].ctypes.data_as(_i16p)
        b.status = self.p["status"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.b = b
        self.pc = None
        self.last_field_factor = ctypes.c_double(-1.0)
        self.cathodes = None   # oracle cathode set, created by define_cathode
        self.n_cathodes = 0

    def __del__(self):
        if getattr(self, "cathodes", None):
            self.lib.oracle_cathodes_destroy(self.cathodes)
            self.cathodes = None

    # ---- imposed solenoid field (src/solenoid.c) ----
    def define_solenoid(self, **namelist):
        d = OSolenoidDesc()
        d.turns = 1
        for key, val in namelist.items():
            setattr(d, key, val)
        for n in ("BzImposed", "BrImposed"):      # allocated by the first solenoid, src/solenoid.c:51-56
            if not getattr(self.f, n):
                setattr(self.f, n, self.arr[n].ctypes.data_as(_dp))
        cur = ctypes.c_double()
        err = ctypes.create_string_buffer(256)
        if self.lib.oracle_define_solenoid(ctypes.byref(self.f), ctypes.byref(d), ctypes.byref(cur), err, 256):
            raise ValueError(err.value.decode())
        return cur.value

    # ---- cathodes (src/cathode.c) ----
    def define_cathode(self, profile=None, **namelist):
        from spiffe_amd import constants as K
        d = OCathodeDesc()
        # defaults of src/cathode.c:32-40
        d.stiffness, d.zoned_emission, d.random_number_seed, d.field_emission_beta = 1.0, 1, 987654321, 1.0
        d.e_mks, d.me_mks, d.k_boltzmann_mks = K.e_mks, K.me_mks, K.k_boltzmann_mks
        for key, val in namelist.items():
            setattr(d, key, val)
        keep = []
        if profile is not None:
            t = np.ascontiguousarray(profile[0], dtype=np.float64)
            a = np.ascontiguousarray(profile[1], dtype=np.float64)
            keep = [t, a]
            d.n_profile = len(t)
            d.profile_time, d.profile_factor = t.ctypes.data_as(_dp), a.ctypes.data_as(_dp)
        if self.cathodes is None:
            self.cathodes = self.lib.oracle_cathodes_create()
        first, count = ctypes.c_int32(), ctypes.c_int32()
        err = ctypes.create_string_buffer(256)
        rc = self.lib.oracle_cathode_define(self.cathodes, ctypes.byref(self.f), ctypes.byref(d), ctypes.byref(first),
                                            ctypes.byref(count), err, 256)
        del keep
        if rc:
            raise ValueError(err.value.decode())
        self.n_cathodes = first.value + count.value
        return list(range(first.value, first.value + count.value))

    def emit_electrons_from_cathode(self, k):
        err = ctypes.create_string_buffer(256)
        n0 = self.b.np
        rc = self.lib.oracle_emit_electrons_from_cathode(self.cathodes, k, ctypes.byref(self.b), ctypes.byref(self.f),
                                                         err, 256)
        if rc:
            raise ValueError(err.value.decode())
        return self.b.np - n0

    def set_current_density(self, k):
        self.lib.oracle_set_current_density(self.cathodes, k, ctypes.byref(self.f))

    def cathode_state(self, k):
        d8 = (ctypes.c_double * 8)()
        i5 = (ctypes.c_int64 * 5)()
        self.lib.oracle_cathode_state(self.cathodes, k, d8, i5)
        names = ["current_density", "ref_current_density", "number_per_step", "electrons_per_macroparticle", "Q",
                 "temperature", "sigma_v", "initial_vz"]
        out = {n: d8[i] for i, n in enumerate(names)}
        out.update(start_step=i5[0], stop_step=i5[1], n_bins=i5[2], step_emission_started=i5[3])
        return out

    # ---- state ----
    def set_field(self, name, a):
        self.arr[name][...] = a

    def field(self, name):
        return self.arr[name].copy()

    def append(self, parts):
        n = len(parts["z"])
        s = self.b.np
        if s + n > self.b.max_np:
            raise ValueError("oracle capacity exceeded")
        for k in PDOUBLES + ["particleID", "generation", "status"]:
            self.p[k][s:s + n] = parts[k]
        self.b.np = s + n
        self.b.npActive += int(np.count_nonzero(np.asarray(parts["status"]) & 1))

    def particles(self):
        n = self.b.np
        return {k: v[:n].copy() for k, v in self.p.items()}

    # ---- the restated functions ----
    def advance_field_solutions(self):
        self.lib.oracle_advance_field_solutions(ctypes.byref(self.f))

    def zero_current_arrays(self):
        self.lib.oracle_zero_current_arrays(ctypes.byref(self.f))

    def multiply_current_arrays(self, factor):
        self.lib.oracle_multiply_current_arrays(ctypes.byref(self.f), factor)

    def add_line_current(self, direction, iposition, istart, iend, J):
        self.lib.oracle_add_line_current(ctypes.byref(self.f), direction, iposition, istart, iend, J)

    def add_particle_currents(self):
        self.lib.oracle_add_particle_currents(ctypes.byref(self.f), ctypes.byref(self.b))

    def interpolate(self, z, r):
        out = [ctypes.c_double() for _ in range(6)]
        self.lib.oracle_interpolate_fields(*[ctypes.byref(o) for o in out], z, r, ctypes.byref(self.f))
        return tuple(o.value for o in out)   # Ez, Er, Ephi, Bz, Br, Bphi

    def advance_particles(self, screens=(), compat=True, z_forces_start=-1.0):
        ns = len(screens)
        scr = (OScreen * max(ns, 1))()
        for k, (z, d) in enumerate(screens):
            scr[k].z_position, scr[k].direction = z, d
        cap_rows = max(ns * max(self.b.np, 1), 1)
        rows = np.zeros(cap_rows, dtype=SCREEN_ROW_DTYPE)
        lost = np.zeros(max(self.b.np, 1), dtype=LOST_ROW_DTYPE)
        nr_, nl_ = ctypes.c_int64(), ctypes.c_int64()
        rc = self.lib.oracle_advance_particles(ctypes.byref(self.f), ctypes.byref(self.b), scr, ns, z_forces_start,
                                               1 if compat else 0,
                                               rows.ctypes.data_as(ctypes.POINTER(OScreenRow)), cap_rows,
                                               ctypes.byref(nr_), lost.ctypes.data_as(ctypes.POINTER(OLostRow)),
                                               len(lost), ctypes.byref(nl_))
        if rc:
            raise RuntimeError("oracle row buffers overflowed")
        return rows[:nr_.value], lost[:nl_.value]

    def setup_poisson_correction(self, step_interval, accuracy=1e-6, maximum_iterations=1000, start_time=0.0,
                                 error_charge_threshold=0.0, guess_code=0):
        pc = OPoisson()
        pc.start_time, pc.step_interval, pc.accuracy = start_time, step_interval, accuracy
        pc.error_charge_threshold, pc.maximum_iterations, pc.guess_code = error_charge_threshold, \
            maximum_iterations, guess_code
        pc.pc_start_step = -1
        self.pc = pc

    def add_poisson_correction(self, space_charge):
        rc = self.lib.oracle_add_poisson_correction(ctypes.byref(self.f), ctypes.byref(self.b),
                                                    1 if space_charge else 0, ctypes.byref(self.pc))
        if rc:
            raise RuntimeError("oracle Poisson correction failed")
        return self.pc.last_iterations

    def solve_poisson_cyl(self, Psi_name, rhs_name, accuracy, maximum_iterations):
        res = ctypes.c_double()
        rhs = self.arr[rhs_name].ctypes.data_as(_dp) if rhs_name else ctypes.cast(None, _dp)
        return self.lib.oracle_solve_poisson_cyl(self.arr[Psi_name].ctypes.data_as(_dp), rhs,
                                                 self.flags.ctypes.data_as(_i16p), self.f.left_bc,
                                                 self.f.right_bc, self.grid.nz, self.grid.nr, self.grid.dz,
                                                 self.grid.dr, accuracy, maximum_iterations, ctypes.byref(res)), \
            res.value

    def compute_imposed_Efield(self, initialize, ramp_time=0.0, flat_top_time=0.0):
        self.f.imposedEFieldRampTime, self.f.imposedEFieldFlatTopTime = ramp_time, flat_top_time
        self.lib.oracle_compute_imposed_efield(ctypes.byref(self.f), 1 if initialize else 0, ctypes.byref(self.pc),
                                               ctypes.byref(self.last_field_factor))

    def step(self, lines=(), screens=(), compat=True, z_forces_start=-1.0, j_filter=0.0):
        """Loop body of src/integrate.c:134-194, emission included, in the order of the device's spf_step.
        j_filter: integrate's J_filter_multiplier m (:110-116): J is scaled by m/(1-m)
        instead of cleared, and by 1-m after the particle and line currents (:168-178)."""
        f = self.f
        f.time = f.time_step * f.dti + f.start_time
        for k in range(self.n_cathodes):
            self.emit_electrons_from_cathode(k)
        if j_filter:
            jm2 = 1 - j_filter
            jm1 = j_filter / jm2
            self.multiply_current_arrays(jm1)
        else:
            self.zero_current_arrays()
        if f.modes & 4:
            self.add_particle_currents()
        for (d, pos, s, e, J) in lines:
            self.add_line_current(d, pos, s, e, J)
        if j_filter:
            self.multiply_current_arrays(jm2)
        if self.pc is not None:
            self.add_poisson_correction((f.modes & 4) != 0)
        self.advance_field_solutions()
        rows, lost = self.advance_particles(screens, compat, z_forces_start)
        for k in range(self.n_cathodes):
            self.set_current_density(k)
        f.time_step += 1
        return rows, lost
