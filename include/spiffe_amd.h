/*
 * spiffe_amd.h -- C ABI of the MI355X-native spiffe time-step library.
 *
 * The reference calls its per-step physics as plain C functions on
 * caller-owned FIELDS/BEAM structs (src/spiffe.h:474-561).  This library owns
 * the field and particle state in HBM instead, behind a context handle; the
 * host keeps the deck and all file I/O.  Every entry
 * point returns 0 on success or a nonzero SPF_E* code; spf_last_error() gives
 * the message.  Calls are stream-ordered on the context's HIP stream and do
 * not block unless documented.  A context is used from one host thread and
 * holds any number of runs; several contexts may work at once from several
 * threads of one process (each run's results are the same bits as when it
 * runs alone), and spf_run_move hands a run from one context to another.
 *
 * Host arrays crossing this boundary use the reference's layout: a 2-D field
 * is (nz+1) x (nr+1) doubles with element [iz][ir] at iz*(nr+1)+ir
 * (src/geometry.c:78-107).  On the device the library stores fields
 * z-contiguous, particles as structure-of-arrays.
 */
#ifndef SPIFFE_AMD_H
#define SPIFFE_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define SPF_OK 0
#define SPF_EINVAL 1      /* bad argument */
#define SPF_EHIP 2        /* HIP runtime error */
#define SPF_ENOMEM 3      /* device allocation failed */
#define SPF_EUNSUPPORTED 4 /* mode not implemented on the device path */
#define SPF_ECAPACITY 5   /* particle or event capacity exceeded */
#define SPF_ENOCONV 6     /* Poisson solve hit its iteration limit (reference returns -1) */

/* ---- mode bits: same values as src/spiffe.h:50-56 ---- */
#define SPF_TM_FIELDS 1u
#define SPF_TE_FIELDS 2u
#define SPF_SPACE_CHARGE 4u
#define SPF_RADIAL_INTERP 8u
#define SPF_LONGIT_INTERP 16u
#define SPF_RADIAL_SMEAR 32u
#define SPF_LONGIT_SMEAR 64u

/* ---- metal flag bits: same values as src/spiffe.h:90-100 ---- */
#define SPF_FL_IS_SURFACE 1
#define SPF_FL_IS_INSIDE 2
#define SPF_FL_EZ_ZERO 4
#define SPF_FL_ER_ZERO 8
#define SPF_FL_IS_VACUUM 16
#define SPF_FL_BPHI_ZERO 32
#define SPF_FL_EPHI_ZERO 64
#define SPF_FL_BZ_ZERO 128
#define SPF_FL_BR_ZERO 256
#define SPF_FL_LAST_ACTIVE_PT 512

#define SPF_DIRICHLET 0
#define SPF_NEUMANN 1

/* particle status bits, src/spiffe.h:219-220 */
#define SPF_PART_ACTIVE 1u
#define SPF_PART_LOST 2u

/* ---- field identifiers for upload/download ---- */
enum spf_field {
  SPF_EZ = 0, SPF_ER, SPF_BPHI, SPF_BPHI2,
  SPF_EPHI, SPF_BR, SPF_BR2, SPF_BZ, SPF_BZ2,
  SPF_JZ, SPF_JR, SPF_JPHI,
  SPF_PSI, SPF_Q,
  SPF_EZ_IMPOSED, SPF_ER_IMPOSED, SPF_PSI_IMPOSED,
  SPF_BZ_IMPOSED, SPF_BR_IMPOSED,
  SPF_N_FIELDS
};

typedef struct spf_ctx spf_ctx;

/* Geometry, modes and run-time options of one run.  Replaces what
 * process_geometry_definition (src/geometry.c:21-369) and the integrate
 * namelist (src/integrate.c:62-108) leave in FIELDS. */
typedef struct {
  int32_t nz, nr;                 /* grid lines, >= 3 each */
  double zmin, zmax, rmax;
  double dz, dr;                  /* (zmax-zmin)/(nz-1), rmax/(nr-1), computed by the host */
  int32_t lower_bc, upper_bc, left_bc, right_bc;
  uint32_t modes;                 /* SPF_* mode bits */
  double dt, start_time;
  double epsilon_o, mu_o, c_mks;  /* physical constants used by the host */
  double QoMC, stiffness;         /* beam charge/(m c) and mass/me */
  int16_t turn_off[6];            /* Ez, Er, Ephi, Bz, Br, Bphi, src/spiffe.h:74-75 */
  double constant_fields[6];      /* Ez, Er, Ephi, Bz, Br, Bphi, src/spiffe.h:118-120 */
  int32_t has_solenoid;           /* nonzero: BzImposed/BrImposed are gathered */
  const int16_t *metal_flags;     /* host (nz+1)*(nr+1), [iz][ir], flags of src/geometry.c:319-362 */
  int64_t max_particles;          /* device particle capacity */
  int32_t n_screens;              /* <= 24 */
  const double *screen_z;         /* screen z positions */
  const int32_t *screen_direction;/* +1 forward, -1 backward */
  int32_t compat;                 /* 1: reproduce compaction and screen-momentum defects of the reference */
  int64_t screen_row_capacity;    /* device ring size for screen rows (0: default) */
  int64_t lost_row_capacity;      /* device ring size for lost rows (0: default) */
  int32_t cell_sort;              /* 0: particles stay in the reference's index order and every sum is
                                     the reference's; N > 0: particles are re-sorted by cell every N steps
                                     and deposition sums in a fixed order of its own (needs compat = 0) */
  int32_t reserved;
} spf_run_desc;

/* Particle SoA on the host, src/spiffe.h:201-221 (status is 32-bit here). */
typedef struct {
  double *Q, *z, *r, *pz, *pr, *r_pphi, *vz, *vr, *vphi;
  double *r0, *z0, *t0, *rmax, *rmin, *zmax, *zmin;
  int32_t *particleID;
  int16_t *generation;
  uint32_t *status;
} spf_particles;

/* One screen-crossing row: the columns of src/screens.c:82-98 plus bookkeeping. */
typedef struct {
  double r, pz, pr, pphi, t, q, r0, t0;
  double rmin, rmax, zmin, zmax;
  int64_t ip;          /* particle index during that step */
  int64_t step;        /* time_step of the row */
  int32_t particleID;
  int16_t generation;
  int16_t screen;
} spf_screen_row;

/* One lost-particle row: the columns of src/advance_particles.c:581-599. */
typedef struct {
  double t;
  float rLost, zLost, pz, pr, r0, z0, q;
  int32_t particleID;
  int64_t ip;
  int64_t step;
  int16_t generation;
  int16_t directionCode;
  int32_t pad;
} spf_lost_row;

/* poisson_correction namelist, src/poisson_correction.nl and defaults at
 * src/poisson_correction.c:37-46 */
typedef struct {
  double start_time;
  int64_t step_interval;
  double accuracy;
  double error_charge_threshold;
  int64_t maximum_iterations;
  int64_t guess_code;   /* 0 none (warm start), 1 line charge, 2 point charge, 3 zero */
} spf_poisson_params;

/* ---- context ---- */
int spf_ctx_create(int32_t device, spf_ctx **out);
int spf_ctx_destroy(spf_ctx *ctx);
const char *spf_last_error(spf_ctx *ctx);
int spf_synchronize(spf_ctx *ctx);
/* HIP stream (hipStream_t) the context launches on, for callers that time or order work around it. */
void *spf_stream(spf_ctx *ctx);

/* ---- run lifecycle; replaces the allocations of src/geometry.c:78-107 ---- */
int spf_run_create(spf_ctx *ctx, const spf_run_desc *desc, int32_t *run_out);
int spf_run_destroy(spf_ctx *ctx, int32_t run);
/* No counterpart in the reference: hands a run (all its device state) from
 * one context to another on the same device, after the run's queued work in
 * `from` has finished; *run_out is its id in `to`.  A serving loop builds runs
 * in a context of its own (uploads, imposed-field solves) while its main
 * context steps the resident ones, then moves them in; the run tables of
 * both contexts are locked for the move, so `to` may be in use by another
 * thread. */
int spf_run_move(spf_ctx *from, int32_t run, spf_ctx *to, int32_t *run_out);
/* bytes of HBM a run with this description would hold (admission control) */
int64_t spf_run_footprint(const spf_run_desc *desc);

/* ---- fields ---- */
int spf_field_upload(spf_ctx *ctx, int32_t run, int32_t field, const double *host);
int spf_field_download(spf_ctx *ctx, int32_t run, int32_t field, double *host);
/* time_step and time = time_step*dt + start_time for the next step, src/integrate.c:138 */
int spf_set_step(spf_ctx *ctx, int32_t run, int64_t time_step);
int spf_get_step(spf_ctx *ctx, int32_t run, int64_t *time_step, double *time);
/* integrate's z_forces_start (src/integrate.nl:19, default -1), which spf_step
 * hands to the push as advance_particles' zForcesStart (src/integrate.c:186-188,
 * src/advance_particles.c:190-200); a run starts with -1 */
int spf_set_z_forces_start(spf_ctx *ctx, int32_t run, double z_forces_start);
/* integrate's J_filter_multiplier m (src/integrate.nl, src/integrate.c:110-116;
 * 0 <= m < 1, default 0): with m != 0 spf_step and spf_step_batch scale J by
 * m/(1-m) instead of clearing it, deposit and add the line currents on top,
 * and scale it by 1-m (src/integrate.c:168-178); a run starts with 0 */
int spf_set_j_filter(spf_ctx *ctx, int32_t run, double multiplier);

/* No counterpart in the reference (which keeps particles in emission order):
 * re-sorts a cell-sort run's particles by cell now (spf_step does it every
 * cell_sort steps).  The particle order is the storage order from then on;
 * particleID names each particle.  SPF_EINVAL when cell_sort is 0. */
int spf_particles_sort(spf_ctx *ctx, int32_t run);
/* No counterpart in the reference: changes a cell-sort run's re-sort interval
 * (> 0; a run created with cell_sort = 0 keeps the reference's order, SPF_EINVAL). */
int spf_set_cell_sort(spf_ctx *ctx, int32_t run, int64_t interval);
/* No counterpart in the reference: the SOR tile shape of this context's solves,
 * 0 (12 half-sweeps per launch on 104x24-point tiles), 1 (8 on 112x32), or -1
 * (the default: 1 from 1.2M grid points solved at once, else 0).  Psi, the
 * iteration counts and every other result are the same for either shape. */
int spf_set_sor_shape(spf_ctx *ctx, int32_t shape);

/* FDTD leapfrog of TM and TE fields; replaces advance_field_solutions, src/integrate.c:277-541 */
int spf_fields_advance(spf_ctx *ctx, int32_t run);
/* J <- 0; replaces zero_current_arrays, src/integrate.c:210-241 */
int spf_currents_reset(spf_ctx *ctx, int32_t run);
/* J <- J*factor; replaces multiply_current_arrays, src/integrate.c:243-275 */
int spf_currents_scale(spf_ctx *ctx, int32_t run, double factor);
/* adds J along one grid line (direction 0: Jz[i][pos], 1: Jr[pos][i], 2: Jphi[i][pos],
 * i = istart..iend); the line loops of add_antenna_excitation, src/antenna.c:137-154 */
int spf_currents_add_line(spf_ctx *ctx, int32_t run, int32_t direction, int64_t iposition,
                          int64_t istart, int64_t iend, double J);
/* NGP current deposition; replaces add_particle_currents, src/space_charge.c:34-295 */
int spf_currents_deposit(spf_ctx *ctx, int32_t run);

/* ---- particles ---- */
/* appends n particles after the current ones (emission sink, src/cathode.c:427-456) */
int spf_particles_append(spf_ctx *ctx, int32_t run, int64_t n, const spf_particles *host);
/* synchronizing: current np and npActive */
int spf_particles_count(spf_ctx *ctx, int32_t run, int64_t *np, int64_t *np_active);
/* synchronizing: copies up to max particles into host arrays (any pointer may be NULL) */
int spf_particles_download(spf_ctx *ctx, int32_t run, const spf_particles *host, int64_t max, int64_t *n);
/* push, gather, loss, screens, compaction; replaces advance_particles,
 * src/advance_particles.c:67-575 (zForcesStart as in src/integrate.c:186-188) */
int spf_particles_push(spf_ctx *ctx, int32_t run, double zForcesStart);
/* synchronizing: copies out and clears the accumulated screen and lost rows */
int spf_events_drain(spf_ctx *ctx, int32_t run, spf_screen_row *rows, int64_t max_rows, int64_t *n_rows,
                     spf_lost_row *lost, int64_t max_lost, int64_t *n_lost);

/* ---- Poisson ---- */
int spf_poisson_setup(spf_ctx *ctx, int32_t run, const spf_poisson_params *p);
/* replaces add_poisson_correction, src/poisson_correction.c:73-454; synchronizing
 * when a solve happens.  *iterations: solve result, -1 not converged, -2 no solve. */
int spf_poisson_correct(spf_ctx *ctx, int32_t run, int32_t space_charge, int64_t *iterations);
/* No counterpart in the reference (which prints the count when verbose,
 * src/poisson_correction.c:279-285): the result of the last correction that
 * solved (-1 not converged, -2 none yet), the number of corrections that
 * solved and the half-sweeps they took in all, for telemetry */
int spf_poisson_stats(spf_ctx *ctx, int32_t run, int64_t *last_iterations, int64_t *solves, int64_t *half_sweeps);
/* replaces solve_poisson_cyl, src/poisson.c:17-130, on the run's Psi (which=0,
 * rhs = Q) or PsiImposed (which=1, no rhs).  Synchronizing. */
int spf_poisson_solve(spf_ctx *ctx, int32_t run, int32_t which, double accuracy, int64_t max_iterations,
                      int64_t *iterations);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The solve of
 * spf_poisson_solve / solve_poisson_cyl (src/poisson.c:17-130) on caller-owned
 * device buffers in the library's layout -- element (iz, ir) at ir*ldz + iz,
 * ldz = nz+1 rounded up to a multiple of 32 -- Psi in/out, rhs (may be NULL)
 * and flags (the run's flag bits) of (nr+1)*ldz elements.  Its scratch memory
 * is one fixed array per GPU inside the library, so it calls no allocator; it
 * holds grids with ldz*(nr+1) <= SPF_NOALLOC_MAX_POINTS and at most
 * SPF_NOALLOC_MAX_TILES SOR tiles, and iteration_limit <=
 * SPF_NOALLOC_MAX_ITERATIONS (larger requests return SPF_EINVAL).  Calls that
 * overlap on one GPU share the array and race.  Synchronizing; *iterations as
 * in spf_poisson_solve. */
#define SPF_NOALLOC_MAX_POINTS (1 << 20)
#define SPF_NOALLOC_MAX_TILES 1024
#define SPF_NOALLOC_MAX_ITERATIONS (1 << 17)
int spf_poisson_solve_noalloc(spf_ctx *ctx, double *Psi, const double *rhs, const int16_t *flags, int32_t nz,
                              int32_t nr, double dz, double dr, int32_t left_bc, int32_t right_bc, double accuracy,
                              int64_t iteration_limit, int64_t *iterations);
/* exactly `count` red-black half-sweeps (half-sweep indices n0 .. n0+count-1 of the
 * omega schedule of src/poisson.c:121-124) with no stopping test: the loop body of
 * solve_poisson_cyl as a fixed-count smoother.  Stream-ordered, not synchronizing. */
int spf_poisson_sweeps(spf_ctx *ctx, int32_t run, int32_t which, int64_t n0, int64_t count);
/* replaces compute_imposed_Efield, src/poisson_correction.c:456-562 */
int spf_imposed_efield(spf_ctx *ctx, int32_t run, int32_t initialize, double ramp_time, double flat_top_time);

/* ---- geometry: boundary file, drawing and metal flags ----
 * What process_geometry_definition (src/geometry.c:21-369) and the drawing
 * routines of src/grid_drawing.c leave in metal_flags, PsiImposed and
 * materialID.  Host-only: no context and no GPU; a nonzero return comes with
 * its message in err (err_len bytes, may be 0). */
typedef struct {
  int32_t nt;                     /* 1 line, 2 arc, 3 start of a new shape, 4 vacuum point */
  double x, y, x0, y0, r;         /* scaled by zr_factor; x, y are DBL_MAX*zr_factor when not given */
  double theta;                   /* end angle of an arc in radians (DBL_MAX when not given) */
  double a, b;                    /* ellipse semi-axes (nt=2, r=0) */
  double potential;
  int16_t material_id, ramp_potential;
} spf_boundary_point;

/* The &po / &point namelists of a boundary file (the text of the file named by
 * define_geometry's `boundary`), src/geometry.c:133-209: per-point defaults,
 * degrees to radians, zr_factor, arc checks and end angles.  With points NULL
 * only the count is returned. */
int spf_geometry_parse(const char *text, double zr_factor, spf_boundary_point *points, int64_t max_points,
                       int64_t *n_points, char *err, int64_t err_len);
/* Draws the boundary, fills vacuum from the axis (or the nt=4 point), trims the
 * grid edges and sets the field flags and FL_LAST_ACTIVE_PT
 * (src/geometry.c:211-362).  flags: (nz+1)*(nr+1) int16 in [iz][ir] order;
 * psi_imposed (boundary potentials) and material may be NULL. */
int spf_geometry_draw(int32_t nz, int32_t nr, double zmin, double zmax, double rmax,
                      const spf_boundary_point *points, int64_t n_points, int16_t *flags, double *psi_imposed,
                      int16_t *material, char *err, int64_t err_len);
/* Adds the field flags and FL_LAST_ACTIVE_PT (src/geometry.c:317-362) to an
 * array holding only FL_IS_SURFACE / FL_IS_INSIDE marks, in place. */
int spf_geometry_flags(int32_t nz, int32_t nr, int16_t *flags);

/* ---- imposed solenoid fields, src/solenoid.c:18-127, src/Bsolenoid.c:17-86 ----
 * The define_solenoid namelist (src/solenoid.nl).  `turns` equispaced current
 * rings from z_start to z_end (mirrored about z = 0 for symmetry +-1) add
 * current x their field to BzImposed/BrImposed at every grid point with
 * ir <= evaluation_radius_limit/dr; Bz_peak rescales the accumulated field to
 * that on-axis maximum; bucking sets the current that cancels the on-axis
 * field already present at z_buck.  Extension (no counterpart in the
 * reference): gaussian != 0 adds instead an on-axis Gaussian of peak Bz_peak
 * centred at gaussian_center with rms gaussian_sigma, and its first-order
 * off-axis expansion Br = -(r/2) dBz/dz.  The run's gather uses the imposed
 * magnetic field from then on.  Synchronizing. */
typedef struct {
  double radius, evaluation_radius_limit, z_start, z_end, current, Bz_peak;
  int64_t turns, symmetry;
  int32_t bucking;
  double z_buck;
  int32_t gaussian;
  double gaussian_center, gaussian_sigma;
} spf_solenoid_desc;

void spf_solenoid_defaults(spf_solenoid_desc *desc);
/* current_used (may be NULL) gets the ring current applied (the bucking one if set) */
int spf_solenoid_add(spf_ctx *ctx, int32_t run, const spf_solenoid_desc *desc, double *current_used);

/* ---- cathode emission, src/cathode.c ----
 * The define_cathode namelist (src/cathode.nl).  spf_cathode_defaults() fills
 * the defaults of src/cathode.c:32-40.  The emission profile arrives as two
 * arrays (the host reads the SDDS file named by `profile`, src/cathode.c:149-166);
 * emission logs are not written.  Random draws come from one stream per run,
 * re-seeded by every define (as random_1(-seed) re-seeds the reference's one
 * stream): Philox4x64-10 keyed by |random_number_seed|, so the particles drawn
 * are the library's own, not the SDDS toolkit's. */
typedef struct {
  double z_position, inner_radius, outer_radius;
  double current_density, temperature, work_function;   /* work_function in eV */
  int32_t field_emission, add_thermal_velocities;
  double field_emission_beta;
  int32_t determine_temperature, autophase;
  double electrons_per_macroparticle, number_per_step;
  double start_time, stop_time, time_offset;
  double initial_pz, initial_omega, stiffness;
  int32_t discretize_radii, spread_over_dt, zoned_emission;
  int32_t distribution_correction_interval;
  int32_t halton_radix_dt, halton_radix_r;
  int64_t random_number_seed;
  int64_t n_profile;                          /* 0: no profile */
  const double *profile_time, *profile_factor;
  double e_mks, me_mks, k_boltzmann_mks;      /* physical constants used by the host */
} spf_cathode_desc;

/* What a cathode holds after its first emission (CATHODE, src/spiffe.h:236-262). */
typedef struct {
  double current_density, ref_current_density, number_per_step, electrons_per_macroparticle;
  double Q, temperature, sigma_v, inner_radius, outer_radius, initial_vz;
  int64_t start_step, stop_step, n_bins, n_emitted, last_count;
} spf_cathode_state;

void spf_cathode_defaults(spf_cathode_desc *desc);
/* process_cathode_definition, src/cathode.c:21-237: checks the namelist and adds
 * one cathode, or one per radial cell with field_emission; returns the index of
 * the first and the number added.  The cathode's q/(m c) and stiffness must equal
 * the run's (src/cathode.c:661-663 copies them into the beam). */
int spf_cathode_define(spf_ctx *ctx, int32_t run, const spf_cathode_desc *desc, int32_t *first, int32_t *count);
/* set_current_density, src/cathode.c:673-715: Richardson-Schottky or field
 * emission from Ez at the cathode centre; synchronizing when a model reads Ez. */
int spf_cathode_set_current_density(spf_ctx *ctx, int32_t run, int32_t cathode);
/* emit_electrons_from_cathode, src/cathode.c:239-671, at the run's current step:
 * the count and the random draws on the host, the new particles written at the
 * end of the device arrays; n_emitted (may be NULL) gets the count. */
int spf_cathode_emit(spf_ctx *ctx, int32_t run, int32_t cathode, int64_t *n_emitted);
int spf_cathode_state_get(spf_ctx *ctx, int32_t run, int32_t cathode, spf_cathode_state *out);
/* the run's nextParticleID (src/spiffe.c:191 starts it at 1): set_next > 0 sets
 * it; next (may be NULL) gets the value in force */
int spf_particle_id_next(spf_ctx *ctx, int32_t run, int64_t set_next, int64_t *next);

/* ---- deck runner ----
 * Runs a spiffe input deck (the command loop of src/spiffe.c:213-296) on the
 * library: define_geometry (boundary file drawn by spf_geometry_*),
 * load_fields, define_antenna, define_cathode, define_solenoid, define_screen,
 * poisson_correction, define_field_saving, set_constant_fields and one
 * integrate.  Screen, lost-particle and field-saving files are written as
 * binary SDDS with the reference's columns; relative file names are taken
 * from the deck's directory, "%s" stands for the rootname.  Diagnostics the
 * runner does not write (field sampling, snapshots, field output, status
 * output) are skipped; commands that would change the physics and are not
 * built (secondary emission, emitters, particle loading, resistors, test
 * beams, on-axis field maps, translation, smoothing) return SPF_EUNSUPPORTED.
 * No counterpart as one function in the reference. */
typedef struct {
  int32_t device;
  int32_t compat;          /* as spf_run_desc.compat (spf_deck_defaults: 1) */
  int32_t verbose;
  int64_t max_particles;   /* particle slots; 0: sized from the cathodes */
  int64_t max_steps;       /* > 0: stop integrate after this many steps */
} spf_deck_options;

typedef struct {
  int64_t steps, particles, emitted, screen_rows, lost_rows, field_saves, max_particles;
} spf_deck_stats;

void spf_deck_defaults(spf_deck_options *opt);
int spf_deck_run(const char *deck_path, const spf_deck_options *opt, spf_deck_stats *stats, char *err,
                 int64_t err_len);

/* ---- SDDS files ----
 * Readers for the SDDS1 files spiffe reads and writes (ASCII or binary pages;
 * double, float, long, short, character and string data).  Host-only. */
int spf_sdds_info(const char *path, int64_t *n_pages, int64_t *n_rows_first_page, char *err, int64_t err_len);
/* numeric column `column` of page `page` (0-based) into out[max]; *n = rows */
int spf_sdds_column(const char *path, int64_t page, const char *column, double *out, int64_t max, int64_t *n,
                    char *err, int64_t err_len);
/* numeric parameter of a page (fixed values included) */
int spf_sdds_parameter(const char *path, int64_t page, const char *name, double *value, char *err,
                       int64_t err_len);

/* ---- phase timers ----
 * No counterpart in the reference: telemetry for benchmarks and serving loops.
 * When enabled, a pair of HIP events on the context's stream brackets one call
 * in every `every` calls of each phase below; a timer accumulates the device
 * time between its events and the number of launches they cover.  Event
 * markers add a few microseconds between kernels, so sparse sampling keeps the
 * measured run close to an untimed one. */
#define SPF_T_STEP 0       /* spf_step */
#define SPF_T_DEPOSIT 1    /* spf_currents_deposit */
#define SPF_T_FDTD 2       /* spf_fields_advance */
#define SPF_T_PUSH 3       /* spf_particles_push: push, compaction and event rows */
#define SPF_T_POISSON 4    /* spf_poisson_correct, when it solves */
#define SPF_T_SOR_BATCH 5  /* SOR batch kernel launches (up to spf_sor_batch_half_sweeps()
                              of the run half-sweeps each), timed per group of back-to-back
                              launches between stopping-rule checks */
#define SPF_T_SORT 6       /* re-sorts of cell-sort runs' particles (spf_particles_sort and in steps) */
#define SPF_T_N 7
/* every > 0 starts timing one call in `every` of each phase and clears the
 * accumulators; 0 stops it */
int spf_timing_enable(spf_ctx *ctx, int32_t every);
/* synchronizing: total milliseconds over the timed launches, the number of timed
 * launches, and (may be NULL) all launches since timing started; arrays of SPF_T_N */
int spf_timing_read(spf_ctx *ctx, double *ms, int64_t *timed, int64_t *launched);
/* half-sweeps one SOR batch launch of this run performs (the library picks
 * one of two build-time tile shapes from the grid size when the run is
 * created), for converting SPF_T_SOR_BATCH times into per-half-sweep rates */
int spf_sor_batch_half_sweeps(spf_ctx *ctx, int32_t run, int32_t *half_sweeps);

/* ---- one whole step of the loop body, src/integrate.c:134-194 ----
 * ramped imposed E (if set); time <- step*dt+start; emission from every
 * cathode of the run; J reset (or the J filter's first scaling); deposit (if
 * SPF_SPACE_CHARGE); n_lines line currents; the J filter's second scaling (if
 * set); Poisson correction (if set up); FDTD; push (with the run's
 * z_forces_start); current densities of the cathodes; step++. */
int spf_step(spf_ctx *ctx, int32_t run, int32_t n_lines, const int32_t *line_dir, const int64_t *line_pos,
             const int64_t *line_start, const int64_t *line_end, const double *line_J);

/* ---- many runs at once ----
 * No counterpart in the reference (one run per process): one whole step
 * (spf_step without line currents) of each of n_runs distinct runs of the
 * context.  The runs' phases are launched together: deposition per run, the
 * Poisson solves of the runs correcting at this step advanced in lockstep by
 * ragged SOR launches over all of them (runs grouped by accuracy and
 * limit), the FDTD of every run in one ragged launch, the push, event rows
 * and compaction of every run in one ragged launch each.  Each run's
 * arithmetic is that of spf_step, so its results are the same bits as when
 * it is stepped alone. */
int spf_step_batch(spf_ctx *ctx, int32_t n_runs, const int32_t *runs);

#ifdef __cplusplus
}
#endif
#endif
