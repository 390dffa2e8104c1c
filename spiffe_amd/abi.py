"""ctypes view of the C ABI declared in include/spiffe_amd.h.

The library is loaded from this package directory (built in-tree by
spiffe_amd.build).  Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspiffe_amd.so")

# error codes
SPF_OK, SPF_EINVAL, SPF_EHIP, SPF_ENOMEM, SPF_EUNSUPPORTED, SPF_ECAPACITY, SPF_ENOCONV = range(7)

# mode bits (src/spiffe.h:50-56)
TM_FIELDS, TE_FIELDS, SPACE_CHARGE, RADIAL_INTERP, LONGIT_INTERP, RADIAL_SMEAR, LONGIT_SMEAR = 1, 2, 4, 8, 16, 32, 64

# metal flag bits (src/spiffe.h:90-100)
FL_IS_SURFACE, FL_IS_INSIDE, FL_EZ_ZERO, FL_ER_ZERO, FL_IS_VACUUM = 1, 2, 4, 8, 16
FL_BPHI_ZERO, FL_EPHI_ZERO, FL_BZ_ZERO, FL_BR_ZERO, FL_LAST_ACTIVE_PT = 32, 64, 128, 256, 512
FL_IS_METAL = FL_IS_SURFACE | FL_IS_INSIDE
DIRICHLET, NEUMANN = 0, 1
PART_ACTIVE, PART_LOST = 1, 2

FIELDS = ["Ez", "Er", "Bphi", "Bphi2", "Ephi", "Br", "Br2", "Bz", "Bz2", "Jz", "Jr", "Jphi", "Psi", "Q",
          "EzImposed", "ErImposed", "PsiImposed", "BzImposed", "BrImposed"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

# phase timers
TIMERS = ["step", "deposit", "fdtd", "push", "poisson", "sor_batch", "cell_sort"]

PARTICLE_DOUBLES = ["Q", "z", "r", "pz", "pr", "r_pphi", "vz", "vr", "vphi",
                    "r0", "z0", "t0", "rmax", "rmin", "zmax", "zmin"]

_dp = ctypes.POINTER(ctypes.c_double)


class RunDesc(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int32), ("nr", ctypes.c_int32),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double), ("rmax", ctypes.c_double),
        ("dz", ctypes.c_double), ("dr", ctypes.c_double),
        ("lower_bc", ctypes.c_int32), ("upper_bc", ctypes.c_int32),
        ("left_bc", ctypes.c_int32), ("right_bc", ctypes.c_int32),
        ("modes", ctypes.c_uint32),
        ("dt", ctypes.c_double), ("start_time", ctypes.c_double),
        ("epsilon_o", ctypes.c_double), ("mu_o", ctypes.c_double), ("c_mks", ctypes.c_double),
        ("QoMC", ctypes.c_double), ("stiffness", ctypes.c_double),
        ("turn_off", ctypes.c_int16 * 6),
        ("constant_fields", ctypes.c_double * 6),
        ("has_solenoid", ctypes.c_int32),
        ("metal_flags", ctypes.POINTER(ctypes.c_int16)),
        ("max_particles", ctypes.c_int64),
        ("n_screens", ctypes.c_int32),
        ("screen_z", _dp),
        ("screen_direction", ctypes.POINTER(ctypes.c_int32)),
        ("compat", ctypes.c_int32),
        ("screen_row_capacity", ctypes.c_int64),
        ("lost_row_capacity", ctypes.c_int64),
        ("cell_sort", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Particles(ctypes.Structure):
    _fields_ = [(n, _dp) for n in PARTICLE_DOUBLES] + [
        ("particleID", ctypes.POINTER(ctypes.c_int32)),
        ("generation", ctypes.POINTER(ctypes.c_int16)),
        ("status", ctypes.POINTER(ctypes.c_uint32)),
    ]


class ScreenRow(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ["r", "pz", "pr", "pphi", "t", "q", "r0", "t0",
                                                "rmin", "rmax", "zmin", "zmax"]] + [
        ("ip", ctypes.c_int64), ("step", ctypes.c_int64),
        ("particleID", ctypes.c_int32), ("generation", ctypes.c_int16), ("screen", ctypes.c_int16)]


class LostRow(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double)] + [(n, ctypes.c_float) for n in
                                           ["rLost", "zLost", "pz", "pr", "r0", "z0", "q"]] + [
        ("particleID", ctypes.c_int32), ("ip", ctypes.c_int64), ("step", ctypes.c_int64),
        ("generation", ctypes.c_int16), ("directionCode", ctypes.c_int16), ("pad", ctypes.c_int32)]


class PoissonParams(ctypes.Structure):
    _fields_ = [("start_time", ctypes.c_double), ("step_interval", ctypes.c_int64),
                ("accuracy", ctypes.c_double), ("error_charge_threshold", ctypes.c_double),
                ("maximum_iterations", ctypes.c_int64), ("guess_code", ctypes.c_int64)]


class BoundaryPoint(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int32)] + [(n, ctypes.c_double) for n in
                                           ["x", "y", "x0", "y0", "r", "theta", "a", "b", "potential"]] + \
               [("material_id", ctypes.c_int16), ("ramp_potential", ctypes.c_int16)]


class SolenoidDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ["radius", "evaluation_radius_limit", "z_start", "z_end", "current",
                                                "Bz_peak"]] + \
               [("turns", ctypes.c_int64), ("symmetry", ctypes.c_int64), ("bucking", ctypes.c_int32),
                ("z_buck", ctypes.c_double), ("gaussian", ctypes.c_int32), ("gaussian_center", ctypes.c_double),
                ("gaussian_sigma", ctypes.c_double)]


class DeckOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("compat", ctypes.c_int32), ("verbose", ctypes.c_int32),
                ("max_particles", ctypes.c_int64), ("max_steps", ctypes.c_int64)]


class DeckStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ["steps", "particles", "emitted", "screen_rows", "lost_rows",
                                               "field_saves", "max_particles"]]


# define_cathode namelist fields in the order of spf_cathode_desc
CATHODE_FIELDS = [
    ("z_position", ctypes.c_double), ("inner_radius", ctypes.c_double), ("outer_radius", ctypes.c_double),
    ("current_density", ctypes.c_double), ("temperature", ctypes.c_double), ("work_function", ctypes.c_double),
    ("field_emission", ctypes.c_int32), ("add_thermal_velocities", ctypes.c_int32),
    ("field_emission_beta", ctypes.c_double),
    ("determine_temperature", ctypes.c_int32), ("autophase", ctypes.c_int32),
    ("electrons_per_macroparticle", ctypes.c_double), ("number_per_step", ctypes.c_double),
    ("start_time", ctypes.c_double), ("stop_time", ctypes.c_double), ("time_offset", ctypes.c_double),
    ("initial_pz", ctypes.c_double), ("initial_omega", ctypes.c_double), ("stiffness", ctypes.c_double),
    ("discretize_radii", ctypes.c_int32), ("spread_over_dt", ctypes.c_int32), ("zoned_emission", ctypes.c_int32),
    ("distribution_correction_interval", ctypes.c_int32),
    ("halton_radix_dt", ctypes.c_int32), ("halton_radix_r", ctypes.c_int32),
    ("random_number_seed", ctypes.c_int64),
    ("n_profile", ctypes.c_int64), ("profile_time", _dp), ("profile_factor", _dp),
    ("e_mks", ctypes.c_double), ("me_mks", ctypes.c_double), ("k_boltzmann_mks", ctypes.c_double),
]


class CathodeDesc(ctypes.Structure):
    _fields_ = CATHODE_FIELDS


class CathodeState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ["current_density", "ref_current_density", "number_per_step",
                                                "electrons_per_macroparticle", "Q", "temperature", "sigma_v",
                                                "inner_radius", "outer_radius", "initial_vz"]] + \
               [(n, ctypes.c_int64) for n in ["start_step", "stop_step", "n_bins", "n_emitted", "last_count"]]


# every entry point declared in include/spiffe_amd.h: name -> (restype, argtypes)
_vp = ctypes.c_void_p
_i32, _i64, _d = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_pi32, _pi64 = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "spf_ctx_create": (ctypes.c_int, [_i32, ctypes.POINTER(_vp)]),
    "spf_ctx_destroy": (ctypes.c_int, [_vp]),
    "spf_last_error": (ctypes.c_char_p, [_vp]),
    "spf_synchronize": (ctypes.c_int, [_vp]),
    "spf_stream": (_vp, [_vp]),
    "spf_run_create": (ctypes.c_int, [_vp, ctypes.POINTER(RunDesc), _pi32]),
    "spf_run_destroy": (ctypes.c_int, [_vp, _i32]),
    "spf_run_move": (ctypes.c_int, [_vp, _i32, _vp, _pi32]),
    "spf_run_footprint": (_i64, [ctypes.POINTER(RunDesc)]),
    "spf_field_upload": (ctypes.c_int, [_vp, _i32, _i32, _dp]),
    "spf_field_download": (ctypes.c_int, [_vp, _i32, _i32, _dp]),
    "spf_set_step": (ctypes.c_int, [_vp, _i32, _i64]),
    "spf_get_step": (ctypes.c_int, [_vp, _i32, _pi64, _dp]),
    "spf_set_z_forces_start": (ctypes.c_int, [_vp, _i32, _d]),
    "spf_set_j_filter": (ctypes.c_int, [_vp, _i32, _d]),
    "spf_fields_advance": (ctypes.c_int, [_vp, _i32]),
    "spf_currents_reset": (ctypes.c_int, [_vp, _i32]),
    "spf_currents_scale": (ctypes.c_int, [_vp, _i32, _d]),
    "spf_currents_add_line": (ctypes.c_int, [_vp, _i32, _i32, _i64, _i64, _i64, _d]),
    "spf_currents_deposit": (ctypes.c_int, [_vp, _i32]),
    "spf_particles_append": (ctypes.c_int, [_vp, _i32, _i64, ctypes.POINTER(Particles)]),
    "spf_particles_count": (ctypes.c_int, [_vp, _i32, _pi64, _pi64]),
    "spf_particles_download": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(Particles), _i64, _pi64]),
    "spf_particles_push": (ctypes.c_int, [_vp, _i32, _d]),
    "spf_particles_sort": (ctypes.c_int, [_vp, _i32]),
    "spf_set_cell_sort": (ctypes.c_int, [_vp, _i32, ctypes.c_int64]),
    "spf_set_sor_shape": (ctypes.c_int, [_vp, _i32]),
    "spf_events_drain": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(ScreenRow), _i64, _pi64,
                                        ctypes.POINTER(LostRow), _i64, _pi64]),
    "spf_poisson_setup": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(PoissonParams)]),
    "spf_poisson_correct": (ctypes.c_int, [_vp, _i32, _i32, _pi64]),
    "spf_poisson_solve": (ctypes.c_int, [_vp, _i32, _i32, _d, _i64, _pi64]),
    "spf_poisson_stats": (ctypes.c_int, [_vp, _i32, _pi64, _pi64, _pi64]),
    "spf_poisson_solve_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _d, _d, _i32, _i32, _d, _i64,
                                                 _pi64]),
    "spf_poisson_sweeps": (ctypes.c_int, [_vp, _i32, _i32, _i64, _i64]),
    "spf_imposed_efield": (ctypes.c_int, [_vp, _i32, _i32, _d, _d]),
    "spf_timing_enable": (ctypes.c_int, [_vp, _i32]),
    "spf_sor_batch_half_sweeps": (ctypes.c_int, [_vp, _i32, _pi32]),
    "spf_timing_read": (ctypes.c_int, [_vp, _dp, _pi64, _pi64]),
    "spf_step": (ctypes.c_int, [_vp, _i32, _i32, _pi32, _pi64, _pi64, _pi64, _dp]),
    "spf_step_batch": (ctypes.c_int, [_vp, _i32, _pi32]),
    "spf_geometry_parse": (ctypes.c_int, [ctypes.c_char_p, _d, ctypes.POINTER(BoundaryPoint), _i64, _pi64,
                                          ctypes.c_char_p, _i64]),
    "spf_geometry_draw": (ctypes.c_int, [_i32, _i32, _d, _d, _d, ctypes.POINTER(BoundaryPoint), _i64,
                                         ctypes.POINTER(ctypes.c_int16), _dp, ctypes.POINTER(ctypes.c_int16),
                                         ctypes.c_char_p, _i64]),
    "spf_geometry_flags": (ctypes.c_int, [_i32, _i32, ctypes.POINTER(ctypes.c_int16)]),
    "spf_solenoid_defaults": (None, [ctypes.POINTER(SolenoidDesc)]),
    "spf_solenoid_add": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(SolenoidDesc), _dp]),
    "spf_deck_defaults": (None, [ctypes.POINTER(DeckOptions)]),
    "spf_deck_run": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(DeckOptions), ctypes.POINTER(DeckStats),
                                    ctypes.c_char_p, _i64]),
    "spf_sdds_info": (ctypes.c_int, [ctypes.c_char_p, _pi64, _pi64, ctypes.c_char_p, _i64]),
    "spf_sdds_column": (ctypes.c_int, [ctypes.c_char_p, _i64, ctypes.c_char_p, _dp, _i64, _pi64, ctypes.c_char_p,
                                       _i64]),
    "spf_sdds_parameter": (ctypes.c_int, [ctypes.c_char_p, _i64, ctypes.c_char_p, _dp, ctypes.c_char_p, _i64]),
    "spf_cathode_defaults": (None, [ctypes.POINTER(CathodeDesc)]),
    "spf_cathode_define": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(CathodeDesc), _pi32, _pi32]),
    "spf_cathode_set_current_density": (ctypes.c_int, [_vp, _i32, _i32]),
    "spf_cathode_emit": (ctypes.c_int, [_vp, _i32, _i32, _pi64]),
    "spf_cathode_state_get": (ctypes.c_int, [_vp, _i32, _i32, ctypes.POINTER(CathodeState)]),
    "spf_particle_id_next": (ctypes.c_int, [_vp, _i32, _i64, _pi64]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libspiffe_amd.so and sets every prototype.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libspiffe_amd.so is not built (%s); run `python -m spiffe_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
