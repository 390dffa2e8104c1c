// spf_emit.h -- host side of cathode emission (src/cathode.c): the namelist
// checks of define_cathode, the current-density models, the per-step count of
// new macroparticles and their random draws.  Plain C++ with no HIP: the
// library's HIP code hands the draws to a device kernel that writes the
// particles into the run (spf_cathode.hip).  Not part of the public ABI.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/spiffe_amd.h"

namespace spf {

// Uniform stream standing in for the SDDS toolkit's random_1: Philox4x64-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11) with key {seed, 0}; block b = 1, 2, ... is Philox(counter {b,0,0,0})
// and its four words are used in order.  A double is (word >> 11) * 2^-53, in
// [0, 1).  gauss() is the Marsaglia polar method, the second deviate of each
// accepted pair kept for the next call (the pairing of gauss_rn).
struct EmitRng {
  uint64_t key0 = 0;
  uint64_t block = 0;       // blocks generated so far
  uint64_t buf[4] = {};
  int pos = 4;              // next unused word of buf; 4 = empty
  bool has_saved = false;
  double saved = 0;
  uint64_t draws = 0;       // words consumed (diagnostics)

  void seed(uint64_t s);
  uint64_t next64();
  double uniform();
  double gauss();
  // gauss_rn_lim(mean, sigma, limit): limit > 0 redraws until |g| <= limit
  double gauss_lim(double mean, double sigma, double limit);
};

void philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
// van der Corput radical inverse of k >= 1 in base b (the Halton points)
double radical_inverse(uint64_t k, int b);

// What emission reads of the run (FIELDS members used by src/cathode.c).
struct CathodeEnv {
  double zmin, zmax, rmax, dz, dr, dti, start_time, time, c_mks;
  int64_t time_step;
  int nz, nr;
};

// CATHODE, src/spiffe.h:236-262.
struct Cathode {
  double z_position = 0, initial_vz = 0, initial_omega = 0;
  double inner_radius = 0, outer_radius = 0;
  double current_density = 0, ref_current_density = 0, temperature = 0, work_function = 0;
  int field_emission = 0;
  double field_emission_beta = 1;
  double start_time = 0, stop_time = 0, time_offset = 0;
  int64_t start_step = -1, stop_step = 0;
  int autophase = 0;
  double number_per_step = 0;
  int64_t step_emission_started = -1;
  double QoMC = 0, Q = 0, electrons_per_macroparticle = 0;
  double inner_r2 = 0, outer_r2 = 0, stiffness = 1;
  int discretize_radii = 0;
  int64_t correction_interval = 0;
  int zoned_emission = 1;
  int64_t n_bins = 0, n_emitted = 0;
  int spread_over_dt = 0;
  std::vector<int64_t> distribution, deficit;
  std::vector<double> area, rmin;
  double totalArea = 0;
  int halton_radix_dt = 0, halton_radix_r = 0;
  uint64_t halton_k_dt = 0, halton_k_r = 0;   // points taken so far
  std::vector<double> t_prof, A_prof;
  double sigmaV = 0;                           // sigmaVMaxwellian
  double e_mks = 0, me_mks = 0, k_boltzmann_mks = 0;
  int64_t last_count = 0;                      // particles emitted by the last call
};

// Per-step output: n particles, in emission order, as five columns of n
// doubles: r, the thermal kicks dvr, dvphi, dvz, and the spread-over-dt time
// dte.  The device writes everything else from these and the cathode.
struct EmitBatch {
  int64_t n = 0;
  bool thermal = false, spread = false;
  int iz_spread = 0;       // z index of the field reads of the spread step
  std::vector<double> cols;
};

// define_cathode, src/cathode.c:21-237: checks the namelist, seeds the stream
// and appends one cathode (or one per radial cell for field emission).
// Returns 0, or nonzero with a message in err.
int cathode_define(const spf_cathode_desc &d, const CathodeEnv &env, EmitRng &rng, std::vector<Cathode> &out,
                   std::string &err);

// set_current_density, src/cathode.c:673-715.  needs_ez() says whether the
// models read the field; ez is constantEz plus Ez + EzImposed at the point
// ez_index() names (with TM fields on).
bool cathode_needs_ez(const Cathode &k);
void cathode_ez_index(const Cathode &k, const CathodeEnv &env, int64_t *iz, int64_t *ir);
void cathode_set_current_density(Cathode &k, double ez);

// One call of emit_electrons_from_cathode, src/cathode.c:239-671: the count
// and the draws of this step, in the reference's order of draws.
int cathode_emit(Cathode &k, const CathodeEnv &env, EmitRng &rng, EmitBatch &out, std::string &err);

// fieldEmissionCurrentDensity, src/cathode.c:799-810
double field_emission_current_density(double E, double beta, double wf);

}  // namespace spf
