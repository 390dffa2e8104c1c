/*
 * spf_oracle.h -- CPU restatement of spiffe's per-time-step 2.5D EM-PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP kernels are
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * Layout follows the reference: every 2-D array is (nz+1) x (nr+1) doubles,
 * element [iz][ir] at offset iz*(nr+1)+ir (src/geometry.c:78-107 allocates
 * them with zarray_2d(..., nz+1, nr+1)).  Arithmetic is FP64 with the
 * reference's expression order; build with -ffp-contract=off.
 */
#ifndef SPF_ORACLE_H
#define SPF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mode bits, src/spiffe.h:50-56 */
#define OFL_TM_FIELDS 1
#define OFL_TE_FIELDS 2
#define OFL_SPACE_CHARGE 4
#define OFL_RADIAL_INTERP 8
#define OFL_LONGIT_INTERP 16
#define OFL_RADIAL_SMEAR 32
#define OFL_LONGIT_SMEAR 64

/* metal flag bits, src/spiffe.h:90-100 */
#define OFL_IS_SURFACE 1
#define OFL_IS_INSIDE 2
#define OFL_EZ_ZERO 4
#define OFL_ER_ZERO 8
#define OFL_IS_VACUUM 16
#define OFL_BPHI_ZERO 32
#define OFL_EPHI_ZERO 64
#define OFL_BZ_ZERO 128
#define OFL_BR_ZERO 256
#define OFL_IS_METAL (OFL_IS_SURFACE | OFL_IS_INSIDE)
#define OFL_LAST_ACTIVE_PT 512

#define O_DIRICHLET 0
#define O_NEUMANN 1

#define OPART_ACTIVE 1u
#define OPART_LOST 2u

/* loss direction codes, src/spiffe.h:604-607 */
#define O_GOING_LEFT 1
#define O_GOING_RIGHT 2
#define O_GOING_UP 3
#define O_GOING_DOWN 4

/* FIELDS, src/spiffe.h:28-131 (the subset the hot path touches) */
typedef struct {
  int32_t nz, nr;
  double zmin, zmax, rmax, dz, dr;
  int32_t lower_bc, upper_bc, right_bc, left_bc;
  double start_time, time, dti;
  int64_t time_step;
  uint32_t modes;
  int16_t turnOffEz, turnOffEr, turnOffEphi, turnOffBz, turnOffBr, turnOffBphi;
  double constantEz, constantEr, constantEphi, constantBz, constantBr, constantBphi;
  double epsilon_o, mu_o, c_mks; /* physical constants supplied by the host */
  double *Ez, *Jz, *Er, *Jr, *Bphi, *Bphi2;
  double *Ephi, *Jphi, *Br, *Br2, *Bz, *Bz2;
  int16_t *metal_flags;
  double *EzImposed, *ErImposed, *PsiImposed;
  double *BzImposed, *BrImposed; /* NULL when no solenoid */
  double imposedEFieldRampTime, imposedEFieldFlatTopTime;
  double *Psi, *Q;
} ofields;

/* BEAM, src/spiffe.h:201-221 */
typedef struct {
  int64_t np, npActive, max_np;
  double QoMC, stiffness;
  double *Q, *z, *r, *pz, *pr, *r_pphi, *vz, *vr, *vphi;
  double *r0, *z0, *t0, *rmax, *rmin, *zmax, *zmin;
  int32_t *particleID;
  int16_t *generation;
  uint32_t *status;
} obeam;

/* SCREEN definition subset, src/spiffe.h:341-352 */
typedef struct {
  double z_position;
  int32_t direction;
  int32_t pad;
} oscreen;

/* one screen row: buffer (r, pz, pr, pphi, t, q, r0, t0) + extents + ids,
 * src/advance_particles.c:382-396 */
typedef struct {
  double r, pz, pr, pphi, t, q, r0, t0;
  double rmin, rmax, zmin, zmax;
  int64_t ip;
  int32_t particleID;
  int16_t generation;
  int16_t screen;
} oscreen_row;

/* one lost-particle row, src/advance_particles.c:471-480 */
typedef struct {
  double t;
  float rLost, zLost, pz, pr, r0, z0, q;
  int32_t particleID;
  int64_t ip;
  int16_t generation;
  int16_t directionCode;
  int32_t pad;
} olost_row;

/* Poisson-correction namelist values, src/poisson_correction.nl */
typedef struct {
  double start_time;
  int64_t step_interval;
  double accuracy;
  double error_charge_threshold;
  int64_t maximum_iterations;
  int64_t guess_code; /* 0 none, 1 line-charge, 2 point-charge, 3 zero */
  int64_t pc_start_step; /* -1 until first use */
  /* outputs of the last call */
  int64_t last_iterations; /* solve result, or -2 if no solve happened */
  double Qreal, Qerror, Qsurface;
} opoisson;

/* define_solenoid namelist, src/solenoid.nl (+ the Gaussian-profile extension) */
typedef struct {
  double radius, evaluation_radius_limit, z_start, z_end, current, Bz_peak;
  int64_t turns, symmetry;
  int32_t bucking;
  double z_buck;
  int32_t gaussian;
  double gaussian_center, gaussian_sigma;
} osolenoid_desc;

void oracle_ellip_ke(double k, double *K, double *E);
void oracle_b_ring(double *B_rho, double *B_z, double rho, double z, double radius);
void oracle_b_solenoid(double *B_rho, double *B_z, double rho, double z, double radius, double zStart, double zEnd,
                       int64_t coils, int64_t symmetry);
int oracle_define_solenoid(ofields *f, const osolenoid_desc *d, double *current_used, char *err, int64_t errlen);

/* define_cathode namelist, src/cathode.nl (the profile file as two arrays) */
typedef struct {
  double z_position, inner_radius, outer_radius;
  double current_density, temperature, work_function;
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
  int64_t n_profile;
  const double *profile_time, *profile_factor;
  double e_mks, me_mks, k_boltzmann_mks;
} ocathode_desc;

/* the uniform stream standing in for random_1 (Philox4x64-10) */
typedef struct {
  uint64_t key0, block;
  uint64_t buf[4];
  int32_t pos, has_saved;
  double saved;
} orng;

typedef struct ocathode_set ocathode_set;

void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
void oracle_rng_seed(orng *g, uint64_t seed);
uint64_t oracle_rng_next64(orng *g);
double oracle_rng_uniform(orng *g);
double oracle_rng_gauss(orng *g);
double oracle_radical_inverse(uint64_t k, int32_t base);
ocathode_set *oracle_cathodes_create(void);
void oracle_cathodes_destroy(ocathode_set *s);
orng *oracle_cathodes_rng(ocathode_set *s);
int64_t oracle_cathodes_next_id(ocathode_set *s);
int oracle_cathode_define(ocathode_set *s, const ofields *f, const ocathode_desc *d, int32_t *first,
                          int32_t *count, char *err, int64_t errlen);
double oracle_field_emission_current_density(double E, double beta, double wf);
void oracle_set_current_density(ocathode_set *s, int32_t ic, const ofields *f);
int oracle_emit_electrons_from_cathode(ocathode_set *s, int32_t ic, obeam *b, const ofields *f, char *err,
                                       int64_t errlen);
/* d8: J, J_ref, number_per_step, electrons_per_macroparticle, Q, T, sigmaV, initial_vz;
 * i5: start_step, stop_step, n_bins, step_emission_started, number of cathodes */
int oracle_cathode_state(ocathode_set *s, int32_t ic, double *d8, int64_t *i5);

void oracle_calculate_flags(int16_t *metal_flags, int32_t nz, int32_t nr);
void oracle_zero_current_arrays(ofields *f);
void oracle_multiply_current_arrays(ofields *f, double factor);
void oracle_advance_field_solutions(ofields *f);
void oracle_add_line_current(ofields *f, int32_t direction, int64_t iposition,
                             int64_t istart, int64_t iend, double J);
double oracle_antenna_current(const double *t, const double *W, int64_t n_pts, int64_t *last_index,
                              double current, double frequency, double phase, double time,
                              int32_t *error);
void oracle_add_particle_currents(ofields *f, const obeam *b);
short oracle_check_loss_code(int64_t iz, int64_t ir, const ofields *f);
void oracle_interpolate_fields(double *Ez, double *Er, double *Ephi, double *Bz, double *Br, double *Bphi,
                               double z, double r, const ofields *f);
void oracle_set_particle_velocities(double *vz, double *vr, double *vphi, double *gamma,
                                    double pz, double pr, double r_pphi, double r, double c_mks);
void oracle_find_loss_coordinates(double *zReturn, double *rReturn, short *directionCode,
                                  double z0, double r0, double z1, double r1, const ofields *f);
int oracle_advance_particles(ofields *f, obeam *b, const oscreen *screens, int32_t n_screens,
                             double zForcesStart, int32_t compat,
                             oscreen_row *rows, int64_t max_rows, int64_t *n_rows,
                             olost_row *lost, int64_t max_lost, int64_t *n_lost);
int64_t oracle_solve_poisson_cyl(double *Psi, const double *driving_term, const int16_t *flags,
                                 int32_t left_bc, int32_t right_bc, int64_t nz, int64_t nr,
                                 double dz, double dr, double accuracy, int64_t iteration_limit,
                                 double *abs_total_residual_out);
void oracle_poisson_omega_schedule(int64_t nz, int64_t nr, double dz, double dr, int64_t n, double *omega);
int oracle_add_poisson_correction(ofields *f, const obeam *b, int32_t space_charge, opoisson *pc);
int oracle_compute_imposed_efield(ofields *f, int32_t initialize, opoisson *pc, double *last_field_factor);
/* J reset + deposit + line currents + Poisson + FDTD + push, src/integrate.c:134-194 */
int oracle_step(ofields *f, obeam *b, const oscreen *screens, int32_t n_screens, int32_t compat,
                opoisson *pc, int32_t n_lines, const int32_t *line_dir, const int64_t *line_pos,
                const int64_t *line_start, const int64_t *line_end, const double *line_J,
                oscreen_row *rows, int64_t max_rows, int64_t *n_rows,
                olost_row *lost, int64_t max_lost, int64_t *n_lost);

#ifdef __cplusplus
}
#endif
#endif
