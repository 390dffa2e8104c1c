// spf_internal.h -- device-side state and kernel launch interfaces shared by
// the HIP translation units of libspiffe_amd.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/spiffe_amd.h"
#include "spf_emit.h"

namespace spf {

constexpr int kMaxScreens = 24;

// Field arrays live z-contiguous: element (iz, ir) at ir*ldz + iz, rows
// ir = 0..nr, ldz >= nz+1 rounded up to 32 doubles (256 B) so each row starts
// on its own cache-line group and a 64-lane wavefront reads 512 contiguous bytes.
__host__ __device__ inline int64_t fix(int ldz, int iz, int ir) { return (int64_t)ir * ldz + iz; }

// Parameters of the FDTD kernel: coefficient tables are computed on the host
// with the reference's expressions so the device only repeats its +,-,*,/ steps.
struct FdtdArgs {
  int nz, nr, ldz;
  int tm, te, left_neu, right_neu, upper_neu;
  double dti, eps0, er_factor, ephi_factor, top_factor, dz, dr;
  const double *Ez, *Er, *Bphi, *Ephi, *Br, *Bz;   // time level n
  double *Ezn, *Ern, *Bphin, *Ephin, *Brn, *Bzn;   // time level n+1
  double *Bphi2, *Br2, *Bz2;                       // half-step B, written in place
  const double *Jz, *Jr, *Jphi;
  const int16_t *flags;
  const int32_t *cut0, *cut1;                      // per-row loop cut-offs
  const double *cp, *cm, *bzf;                     // per-row coefficients
};

struct PushArgs {
  int nz, nr, ldz;
  int tm, te, longit, radial, has_sol, imp_e;
  double zmin, zmax, rmax, dz, dr, dt, time, c, QoMC, stiffness, zForcesStart;
  const double *Ez, *Er, *Ephi, *Bphi2, *Br2, *Bz2, *EzImp, *ErImp, *BzImp, *BrImp;
  const int16_t *flags;
  const uint8_t *open;              // open cells: no loss-test or gather flag within reach
  int16_t turnoff[6];
  double cst[6];
  double *Q, *z, *r, *pz, *pr, *rpphi, *vz, *vr, *vphi, *r0, *z0, *t0, *rmax_p, *rmin_p, *zmax_p, *zmin_p;
  uint32_t *status;
  const int64_t *np;
  // per-step event staging (indexed by particle)
  uint32_t *evt;
  uint8_t *evflag;
  int32_t *pcount;                  // events per push workgroup
  unsigned long long *ev_total;     // events of the step, summed over the workgroups that have any
  double *ev_zold, *ev_rold, *ev_vz, *ev_vr, *ev_pz0, *ev_pr0, *ev_rp0;
  double *last_old;                 // old pz, pr, r_pphi of particle np-1
  unsigned long long *new_active;   // activations this step
  int nscreens;
  double sz[kMaxScreens];
  int sdir[kMaxScreens];
};

// Event rows of one step (single workgroup kernel) and compaction.
struct RowsArgs {
  int nz, nr, ldz, compat, nscreens;
  double zmin, zmax, dz, dr, dt, time;
  int64_t step;
  const int16_t *flags;
  const uint8_t *evflag;            // per particle: has an event this step
  const int32_t *pcount;            // events per push workgroup
  unsigned long long *ev_total;     // events of the step; cleared for the next step
  int64_t npb, n_ub;                // push workgroups, particle slots they covered
  int32_t *evlist;                  // built here: particles with events, in index order
  int64_t *n_ev;
  int64_t *np, *np_active;
  unsigned long long *new_active;   // read, then cleared for the next step
  const uint32_t *evt;
  const double *ev_zold, *ev_rold, *ev_vz, *ev_vr, *ev_pz0, *ev_pr0, *ev_rp0, *last_old;
  double *Q, *z, *r, *pz, *pr, *rpphi, *vz, *vr, *vphi, *r0, *z0, *t0, *rmax_p, *rmin_p, *zmax_p, *zmin_p;
  int32_t *pid;
  int16_t *gen;
  uint32_t *status;
  double sz[kMaxScreens];
  int sdir[kMaxScreens];
  spf_screen_row *srows;
  int64_t *n_srows;
  int64_t srow_cap;
  spf_lost_row *lrows;
  int64_t *n_lrows;
  int64_t lrow_cap;
  int32_t *move_src, *move_dst;     // scratch, capacity = particle capacity
  uint8_t *tflag;                   // scratch bytes, capacity >= particle capacity
  int32_t *overflow;
  // the multi-workgroup rows pipeline of large runs (rows_pipeline_for)
  int32_t *evoff;                   // first event of each push workgroup
  int32_t *tcnt, *toff;             // per row category (screens, then lost) and event tile: count, offset
  int64_t ntmax;                    // tile stride of tcnt/toff (event tiles of the largest step)
  int64_t *cbase;                   // per category: first row; [nscreens+1] lost count, [nscreens+2] moves
};

struct Run {
  bool live = false;
  spf_run_desc d{};
  int ldz = 0;
  int64_t cells = 0;        // (nr+1)*ldz
  int64_t time_step = 0;
  double time = 0;          // time of the current step, src/integrate.c:138
  // fields: ping-pong pairs for E and B, single arrays for the rest
  double *E2[6][2] = {};    // Ez, Er, Bphi, Ephi, Br, Bz
  int cur[6] = {};
  double *Bphi2 = nullptr, *Br2 = nullptr, *Bz2 = nullptr;
  double *Jz = nullptr, *Jr = nullptr, *Jphi = nullptr;
  double *Psi = nullptr, *Q = nullptr;
  double *PsiSpare[5] = {};   // spare buffers the SOR batches rotate through
  double *EzImp = nullptr, *ErImp = nullptr, *PsiImp = nullptr, *BzImp = nullptr, *BrImp = nullptr;
  int16_t *flags = nullptr;
  uint8_t *open = nullptr;      // per cell: the push may skip its flag tests there
  int32_t *cut0 = nullptr, *cut1 = nullptr;
  double *cp = nullptr, *cm = nullptr, *bzf = nullptr;
  double *sor_c = nullptr, *sor_d = nullptr, *volume = nullptr;
  double sor_c_host0 = 0;   // sor_c[0] on the host (axis-row reciprocal)
  // particles
  int64_t cap = 0;
  int64_t np_ub = 0;        // host-side upper bound of the device particle count
  double *P[16] = {};       // Q z r pz pr r_pphi vz vr vphi r0 z0 t0 rmax rmin zmax zmin
  int32_t *pid = nullptr;
  int16_t *gen = nullptr;
  uint32_t *status = nullptr;
  // cell-sort runs: the particle set the re-sort gathers into (swapped with
  // the current one afterwards), and the reduce deposition's partial counts
  int64_t cell_sort = 0;    // re-sort interval in steps (0: reference order kept)
  double *P2[16] = {};
  int32_t *pid2 = nullptr;
  int16_t *gen2 = nullptr;
  uint32_t *status2 = nullptr;
  int32_t *wg_count = nullptr;
  int64_t *wg_off = nullptr;
  // [0] np, [1] npActive, [2] n_ev, [3] n_srows, [4] n_lrows, [5] events of a
  // push, [7] newly active particles of a push, [8] event-ring overflow flag,
  // [9] long segments of a deposit, [10] partials of a reduce deposition; [6] unused
  int64_t *counters = nullptr;
  int64_t *krange = nullptr;    // deposition key range: 256 min and 256 max slots, then the range (2)
  unsigned long long *new_active = nullptr;
  int32_t *overflow = nullptr;
  // per-step events
  uint32_t *evt = nullptr;
  uint8_t *evflag = nullptr;
  double *ev[7] = {};
  double *last_old = nullptr;
  int32_t *evlist = nullptr;
  int32_t *move_src = nullptr, *move_dst = nullptr;
  spf_screen_row *srows = nullptr;
  spf_lost_row *lrows = nullptr;
  int64_t srow_cap = 0, lrow_cap = 0;
  // ordered scatter-add (deposition and charge removal): up to 3 components per particle
  uint32_t *keys = nullptr, *skeys = nullptr;
  double *cz = nullptr, *scz = nullptr;   // contributions, unsorted and sorted
  int32_t *vals = nullptr;                // list of long segments
  int32_t *pcount = nullptr;              // events per push workgroup
  int32_t *evoff = nullptr;               // rows pipeline: first event of each push workgroup
  int32_t *tcnt = nullptr, *toff = nullptr;   // rows pipeline: per category and event tile
  int64_t *cbase = nullptr;               // rows pipeline: category bases and counts
  bool imposed_e = false;                 // EzImposed/ErImposed may hold nonzero values
  int32_t *long_list = nullptr;           // long segments, folded one per wavefront
  uint8_t *gflag = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // staging
  double *stage = nullptr;      // (nz+1)*(nr+1) doubles
  void *append_stage = nullptr;
  size_t append_cap = 0;
  double **Pdev = nullptr;      // device copy of P[16]
  // Poisson
  bool pc_set = false;
  spf_poisson_params pc{};
  int64_t pc_start_step = -1;
  double *omega = nullptr;      // omega per half-sweep
  int64_t omega_n = 0;
  double *red = nullptr;        // reduction partials
  int64_t red_n = 0;
  int32_t *red_counter = nullptr;
  int64_t *sor_state = nullptr; // [0] done, [1] iterations, [2] stop left to the host (within rounding)
  int sor_shape = 0;            // SOR tile shape of the current or last solve (buffers are sized for shape 0)
  uint32_t *sor_masks = nullptr; // per-tile lane masks of the SOR batches
  double *sor_sums = nullptr;   // [0] drive sum, [1] last residual sum, [2..] misc
  double last_field_factor = -1;
  int64_t pc_last_iters = -2;   // result of the last correction (-2: none solved)
  int64_t pc_solves = 0, pc_half_sweeps = 0;   // corrections solved and half-sweeps they took
  std::vector<double> screen_z;
  std::vector<int> screen_dir;
  double ramp_time = 0, flat_top_time = 0;
  double z_forces_start = -1;   // integrate's z_forces_start (src/integrate.nl:19), used by spf_step
  // integrate's J_filter_multiplier m (src/integrate.c:110-116): when nonzero
  // the steps scale J by jm1 = m/(1-m) instead of clearing it and by jm2 = 1-m
  // after the particle and line currents
  double jfm = 0, jm1 = 1, jm2 = 0;
  // cathode emission: host-side state, one uniform stream per run
  EmitRng emit_rng;
  std::vector<Cathode> cathodes;
  int64_t next_pid = 1;         // nextParticleID, src/spiffe.c:191
  double *emit_stage = nullptr; // per-particle draws of one emission call
  size_t emit_cap = 0;
};

// ---- launchers implemented in the .hip files ----
hipError_t launch_fdtd(const FdtdArgs &a, hipStream_t s);
// ragged FDTD step over several runs: table into `host` (fdtd_multi_table_bytes),
// copied to the device, then one tile launch and one edge launch
size_t fdtd_multi_table_bytes(int nruns);
void fdtd_multi_table(const FdtdArgs *args, int nruns, void *host, int64_t *tile_blocks, int64_t *edge_blocks);
hipError_t launch_fdtd_multi(const void *dev_table, int nruns, int64_t tile_blocks, int64_t edge_blocks,
                             hipStream_t s);
hipError_t launch_transpose_in(const double *host_layout, double *dev, int nz, int nr, int ldz, hipStream_t s);
hipError_t launch_transpose_out(const double *dev, double *host_layout, int nz, int nr, int ldz, hipStream_t s);
hipError_t launch_transpose_in_i16(const int16_t *host_layout, int16_t *dev, int nz, int nr, int ldz, hipStream_t s);
hipError_t launch_fill3(double *p0, double *p1, double *p2, int64_t n, double v, hipStream_t s);
hipError_t launch_set_i64(int64_t *p, int64_t v, hipStream_t s);
hipError_t launch_scale(double *p, int64_t n, double f, hipStream_t s);
hipError_t launch_add_line(double *J, int ldz, int direction, int64_t ipos, int64_t istart, int64_t iend, double v,
                           hipStream_t s);
hipError_t launch_push(const PushArgs &a, int64_t n_ub, hipStream_t s);
hipError_t launch_open_cells(const int16_t *flags, int ldz, int nz, int nr, uint8_t *open, hipStream_t s);
hipError_t launch_rows(const RowsArgs &a, hipStream_t s);
// events per tile of the rows pipeline, and the particle slots from which a
// run's rows go through it (below, one workgroup builds them: fewer launches)
constexpr int kRowsTile = 1024;
constexpr int64_t kRowsPipelineMin = 1 << 20;
inline int64_t rows_tiles(int64_t cap) { return (cap + kRowsTile - 1) / kRowsTile + 1; }
// ragged push of several runs (table from push_multi_table, push_multi_table_bytes
// long, copied to the device): one push launch, then one rows workgroup per run
size_t push_multi_table_bytes(int nruns);
int64_t push_multi_table(const PushArgs *pa, const int64_t *n_ub, const RowsArgs *ra, int nruns, void *host);
hipError_t launch_push_multi(const void *dev_table, int nruns, int64_t total_blocks, hipStream_t s);
size_t append_stage_bytes(int64_t n);
hipError_t launch_append_kernels(Run &R, const void *stage, int64_t n, int64_t n_active, double *const *Pdev,
                                 hipStream_t s);
// np += n, npActive += n_active after n particles were written past np
hipError_t launch_append_commit(Run &R, int64_t n, int64_t n_active, hipStream_t s);
// cathode emission (spf_cathode.hip): one cathode's emission at the current step,
// and its current-density update
int cathode_emit_step(spf_ctx *c, int32_t run, int32_t cathode, int64_t *n_emitted);
int cathode_current_density(spf_ctx *c, int32_t run, int32_t cathode);
// deposition
// Jz, Jr, Jphi += particle currents; with clear_first they are set to +0 first
hipError_t deposit_currents(Run &R, hipStream_t s, bool clear_first = false);
// J reset and deposition of n cell-sort runs (space charge, TM or TE fields,
// particles present) in one set of launches, each run the bits of its own
// reduce deposition; hipErrorInvalidValue when their key spaces do not fit 32 bits
hipError_t deposit_currents_multi(spf_ctx *c, Run *const *runs, int n);
// cell-sort runs: stable re-sort of the particles by half-cell (gathered into
// the second particle set, which then becomes the current one)
hipError_t sort_particles(Run &R, hipStream_t s);
// workgroups of the reduce deposition (one per 1024 particle slots)
inline int64_t reduce_groups(int64_t n_ub) { return (n_ub + 1023) / 1024; }
int deposit_cell_bits(int nz, int nr);
// key slots per particle and component of the ordered deposition: 4 with
// area-weighted (smeared) deposition, else 1
int deposit_slots_per_particle(uint32_t modes);
// stable radix sort of n (key, value) pairs on the low key_bits bits of the
// keys; the input (k0, v0) and (k1, v1) alternate as pass input and output,
// *kres/*vres name the pair holding the result
size_t radix_sort_tmp_bytes(int64_t n);
hipError_t radix_sort_pairs(uint32_t *k0, double *v0, uint32_t *k1, double *v1, int64_t n, int key_bits, void *tmp,
                            size_t tmp_bytes, uint32_t **kres, double **vres, hipStream_t s);
// The same sort over keys whose values other than `discard` lie in the device
// range [range[0], range[1]] (uint64; range[1] < range[0]: none): 7-bit digits
// of the key's offset in the range (discard after every real key), and only
// the passes the range needs do work.  The result is in (k0, v0) when an even
// number of passes worked (ranged_sort_in_k1 on the device), else in (k1, v1).
inline int sorted_passes_max(int key_bits) { return (key_bits + 6) / 7; }
// n_dev (optional): the number of pairs, read on the device (at most n); the
// tables keep n's layout and only the tiles holding pairs are worked on
hipError_t radix_sort_pairs_ranged(uint32_t *k0, double *v0, uint32_t *k1, double *v1, int64_t n, int key_bits,
                                   const uint64_t *range, uint32_t discard, void *tmp, size_t tmp_bytes,
                                   hipStream_t s, const int64_t *n_dev = nullptr);
__device__ inline bool ranged_sort_in_k1(const uint64_t *range) {
  const uint64_t lo = range[0], hi = range[1];
  const uint32_t span = hi >= lo ? (uint32_t)(hi - lo) + 1u : 0u;
  const int nbits = span ? 32 - __builtin_clz(span) : 0;
  return ((nbits + 6) / 7) & 1;
}
// Poisson
hipError_t launch_gauss_q(Run &R, hipStream_t s);
// Q -= particle charge (src/poisson_correction.c:184-243); with qreal_sums the
// fixed-order sum of the removed charges' negatives and of their magnitudes
// go to qreal_sums[0], [1]
hipError_t launch_q_subtract(Run &R, hipStream_t s, double *qreal_sums = nullptr);
// Qreal summed in particle order, one rounded addition at a time, as the reference does
hipError_t launch_qreal_serial(Run &R, double *out, hipStream_t s);
// fixed-order plain and absolute sums (out[0], out[1]) of a region or a vector;
// partials needs 1024 entries
hipError_t launch_sum2_region(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn,
                              double *out, hipStream_t s);
hipError_t launch_sum2_vec(const double *a, int64_t n, double *partials, int64_t pn, double *out, hipStream_t s);
// serial sums in the reference's loop orders: iz outer / ir inner (absval: of |a|), and ir outer / iz inner
hipError_t launch_serial_sum_iz_ir(const double *a, int ldz, int iz_n, int ir_n, int absval, double *out,
                                   hipStream_t s);
hipError_t launch_serial_sum_ir_iz(const double *a, int ldz, int iz_n, int ir_n, double *out, hipStream_t s);
// sum over ir < ir_n (outer), iz < iz_n of a[iz][ir] * (iz*dz + zmin), serially in that order
hipError_t launch_serial_moment_ir_iz(const double *a, int ldz, int iz_n, int ir_n, double dz, double zmin, double *out,
                                      hipStream_t s);
// bound on how far sums of n terms with magnitude sum `mag` can differ between two summation orders
double order_margin(int64_t n, double mag);
hipError_t launch_q_to_rhs(Run &R, hipStream_t s);
hipError_t launch_sum_abs(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn, double *out,
                          hipStream_t s);
hipError_t launch_sum_plain(const double *a, int ldz, int iz_n, int ir_n, double *partials, int64_t pn, double *out,
                            hipStream_t s);
hipError_t launch_sor_reset(Run &R, hipStream_t s);
int sor_shape_for(int64_t points);   // SOR tile shape of a solve over this many grid points (0 small, 1 large)
int sor_max_batch();                 // the most half-sweeps a batch launch performs, over both shapes
int64_t sor_tiles(const Run &R);
int sor_check_group(const Run &R);   // batches per stopping-rule check (<= 4)
int64_t sor_mask_words(const Run &R);
hipError_t launch_sor_masks(Run &R, hipStream_t s);   // before the batches of a solve
int sor_batch(const Run &R);         // half-sweeps per batch launch of the run's shape

// check_k > 0: the launch carries one extra workgroup that applies the
// stopping rule to the previous check group (check_k half-sweeps from
// check_n0, partial rows from check_row) while the tiles run
hipError_t launch_sor_batch(Run &R, const double *Pin, double *Pout, const double *rhs, int64_t n0, int k,
                            int partial_row, int force, hipStream_t s, int64_t check_row = 0, int64_t check_n0 = 0,
                            int check_k = 0, double accuracy = 0);
hipError_t launch_sor_check(Run &R, int64_t check_row, int64_t n0, int k, double accuracy, hipStream_t s);
// Ragged SOR batch launches over several runs solving in lockstep (same
// half-sweep range, each run its own grid, ring of 6 Psi buffers, partial
// rows and state).  sor_multi_table writes the per-run table into `host`
// (sor_multi_table_bytes) for a copy to `dev`, and returns the launch's
// workgroup count; launch_sor_multi runs batch index b (ring[b%6] -> ring[(b+1)%6]).
size_t sor_multi_table_bytes(int nruns);
int64_t sor_multi_table(Run *const *runs, const double *const *rhs, double *const (*rings)[6], int nruns,
                        double accuracy, void *host, void *dev, const int32_t **first_dev);
hipError_t launch_sor_multi(int shape, const void *dev_table, const int32_t *first_dev, int nruns,
                            int64_t total_blocks, int64_t n0, int64_t b, int k, int64_t partial_row, int64_t check_row,
                            int64_t check_n0, int check_k, hipStream_t s);   // every run of the table has `shape`
// half-sweep n in place with the reference's expressions, |residual| per point to dump[ir*nz+iz]
hipError_t launch_sor_exact_sweep(Run &R, double *P, const double *rhs, int64_t n, double *dump, hipStream_t s);
// done flag, iteration count and the host-decision flag back to "running"
hipError_t launch_sor_state_clear(Run &R, hipStream_t s);
hipError_t launch_field_from_psi(Run &R, int imposed, double field_factor, hipStream_t s);
hipError_t launch_zero_psi(Run &R, hipStream_t s);
// the guesses of src/poisson_correction.c:564-615 for iz < nz, ir < nr: Psi
// row ir set to row[ir] (line charge, the rows formed on the host), or the
// point-charge potential of charge *q at z = *centroid
hipError_t launch_line_guess(Run &R, const double *row, hipStream_t s);
hipError_t launch_point_guess(Run &R, const double *q, const double *centroid, hipStream_t s);
// resolves, once per device, the address of the no-allocation solve's scratch array (spf_noalloc.hip)
int noalloc_setup(int device);

}  // namespace spf

// The context behind the opaque spf_ctx handle.
struct spf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<spf::Run *> runs;
  std::mutex runs_mu;   // guards `runs`: spf_run_move may insert from another thread
  std::string err;
  int64_t *pinned = nullptr;   // small readbacks
  void *hstage = nullptr;      // pinned staging for appends
  size_t hstage_bytes = 0;
  hipEvent_t stage_done = nullptr;
  hipEvent_t sor_ev[2] = {nullptr, nullptr};   // ends of the two SOR check windows in flight
  // phase timers (SPF_T_*): event pairs recorded on the stream, resolved lazily
  int64_t timing_every = 0;    // 0: off; N: one call in N of each phase is timed
  int64_t t_calls[SPF_T_N] = {};
  int64_t t_launched[SPF_T_N] = {};   // launches of each phase while timing, sampled or not
  struct Pending {
    int id;
    int64_t n;                 // launches covered by the event pair
    hipEvent_t e0, e1;
  };
  std::vector<hipEvent_t> ev_free;
  std::vector<Pending> ev_pending;
  double t_ms[SPF_T_N] = {};
  int64_t t_count[SPF_T_N] = {};
  // ragged (multi-run) launches: argument tables staged in pinned slots and
  // copied to device slots, round-robin; a slot's host side is rewritten only
  // after its copy has completed (event per slot)
  int sor_shape = -1;           // SOR tile shape of this context's solves (-1: by the points solved)
  // batched reduce deposition: packed partials and their sort (grown on demand)
  int64_t dep_cap = 0, dep_runs = 0;
  uint32_t *dep_k0 = nullptr, *dep_k1 = nullptr;
  double *dep_v0 = nullptr, *dep_v1 = nullptr;
  void *dep_tmp = nullptr;
  size_t dep_tmp_bytes = 0;
  int32_t *dep_long = nullptr;
  int64_t *dep_krange = nullptr, *dep_misc = nullptr, *dep_run_off = nullptr;
  static constexpr int kTableSlots = 16;
  char *bt_host = nullptr, *bt_dev = nullptr;
  size_t bt_slot = 0;
  int bt_next = 0, bt_cur = 0;
  hipEvent_t bt_ev[kTableSlots] = {};
  bool bt_used[kTableSlots] = {};
  int64_t *bw_flags = nullptr;   // pinned readback of per-run flags
  size_t bw_flags_n = 0;
};

namespace spf {

inline int fail(spf_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

// Brackets one phase (covering `launches` launches) with a pair of HIP events
// on the context's stream when timing is enabled and this call is sampled;
// the elapsed device time is added to timer `id` later.
struct PhaseTimer {
  spf_ctx *c;
  int id;
  int64_t n;
  hipEvent_t e0 = nullptr;
  PhaseTimer(spf_ctx *ctx, int timer_id, int64_t launches = 1);
  ~PhaseTimer();
  PhaseTimer(const PhaseTimer &) = delete;
  PhaseTimer &operator=(const PhaseTimer &) = delete;
};

void resolve_timers(spf_ctx *c, bool wait);

// arguments of one run's FDTD step and push (as spf_fields_advance and
// spf_particles_push launch them); fdtd_flip swaps the E/B buffers after the step
void fdtd_args(Run *R, FdtdArgs &a);
void fdtd_flip(Run *R);
void push_args(Run *R, double zForcesStart, PushArgs &a, RowsArgs &w);

// a table slot of at least `bytes` for a ragged launch: host (pinned) and
// device addresses; fill the host side, then batch_table_upload it
int batch_table(spf_ctx *c, size_t bytes, void **host, void **dev);
hipError_t batch_table_upload(spf_ctx *c, const void *host, void *dev, size_t bytes);
// the batched deposition's scratch for `pairs` packed pairs and `runs` runs
int deposit_scratch(spf_ctx *c, int64_t pairs, int64_t runs);
// c->bw_flags holds at least n int64
int batch_flags(spf_ctx *c, size_t n);

inline Run *get_run(spf_ctx *c, int32_t run) {
  if (!c) return nullptr;
  std::lock_guard<std::mutex> lock(c->runs_mu);
  if (run < 0 || run >= (int32_t)c->runs.size() || !c->runs[run] || !c->runs[run]->live) return nullptr;
  return c->runs[run];
}

// stores R in a free slot of c's run table and returns its id
inline int32_t put_run(spf_ctx *c, Run *R) {
  std::lock_guard<std::mutex> lock(c->runs_mu);
  for (size_t k = 0; k < c->runs.size(); k++)
    if (!c->runs[k]) {
      c->runs[k] = R;
      return (int32_t)k;
    }
  c->runs.push_back(R);
  return (int32_t)c->runs.size() - 1;
}

inline void clear_run(spf_ctx *c, int32_t run) {
  std::lock_guard<std::mutex> lock(c->runs_mu);
  c->runs[run] = nullptr;
}

}  // namespace spf

#define HIPCALL(ctx, expr)                                                                         \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return spf::fail(ctx, e_ == hipErrorOutOfMemory ? SPF_ENOMEM : SPF_EHIP,                     \
                       std::string(#expr) + ": " + hipGetErrorString(e_));                         \
  } while (0)
