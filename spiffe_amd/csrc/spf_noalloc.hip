// spf_noalloc.hip -- spf_poisson_solve_noalloc: the red-black SOR solve of
// src/poisson.c:17-130 on caller-owned device buffers, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Everything the solve needs beyond
// its inputs and outputs -- the second Psi buffer the batches ping-pong with,
// the residual dump of an exact stopping decision, the lanes' masks, the
// per-tile partial sums, the omega table, the row coefficients and the
// solve's state -- lives in one file-scope device array, `scratch`, of fixed
// size.  Each GPU holds its own copy until the process ends; its address is
// resolved once per device when a context is created on it.  Every call on a
// GPU uses that GPU's one array, so calls that overlap on one GPU (from two
// threads or two contexts) race; each call leaves the array holding its own
// state and re-initialises what must start afresh with kernels (no
// hipMemsetAsync, so the call can sit inside a captured graph).
#include <cmath>
#include <cstring>
#include <mutex>

#include "spf_internal.h"

using namespace spf;

namespace {

// limits of the array (see include/spiffe_amd.h)
constexpr int64_t kMaxPoints = SPF_NOALLOC_MAX_POINTS;   // ldz * (nr+1)
constexpr int64_t kMaxTiles = SPF_NOALLOC_MAX_TILES;     // SOR tiles of the grid
constexpr int64_t kMaxIter = SPF_NOALLOC_MAX_ITERATIONS; // iteration_limit
constexpr int64_t kTileWords = 512;                      // lane-mask words per tile (>= threads per SOR tile)
constexpr int64_t kPartialRows = 16;                     // >= half-sweeps per batch
constexpr int64_t kRows = 8192;                          // >= nr + 1
constexpr int kMaxDevices = 64;

// layout in doubles
constexpr int64_t OFF_SPARE = 0;
constexpr int64_t OFF_DUMP = OFF_SPARE + kMaxPoints;
constexpr int64_t OFF_MASKS = OFF_DUMP + kMaxPoints;                    // uint32 words, two per double
constexpr int64_t OFF_PARTIALS = OFF_MASKS + kMaxTiles * kTileWords / 2;
constexpr int64_t OFF_OMEGA = OFF_PARTIALS + kPartialRows * kMaxTiles;
constexpr int64_t OFF_C = OFF_OMEGA + kMaxIter;
constexpr int64_t OFF_D = OFF_C + kRows;
constexpr int64_t OFF_SUMS = OFF_D + kRows;                             // 16 doubles
constexpr int64_t OFF_STATE = OFF_SUMS + 16;                            // 4 int64
constexpr int64_t OFF_COUNTER = OFF_STATE + 4;                          // 1 int32
constexpr int64_t kScratchDoubles = OFF_COUNTER + 2;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};          // device address of `scratch` on each device
double g_omega_rj[kMaxDevices] = {};      // r_jacobi and length of the omega table the array holds
int64_t g_omega_n[kMaxDevices] = {};

// omega for half-sweeps 0 .. n-1 by the recurrence of src/poisson.c:121-124
// (one thread: each value depends on the previous one; only +, -, *, / on
// the host-computed r_jacobi, so the bits are the host's), and the row
// coefficients c, d of src/poisson.c:58-66
__global__ void sor_tables_kernel(double r_jacobi, int64_t n_omega, double *omega, int nr, double dr, double *c,
                                  double *d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double om = 1.0;
    for (int64_t k = 0; k < n_omega; k++) {
      omega[k] = om;
      if (k == 0)
        om = 1.0 / (1 - r_jacobi * r_jacobi / 2);
      else
        om = 1.0 / (1 - r_jacobi * r_jacobi * om / 4);
    }
  }
  for (int ir = (int)(blockIdx.x * blockDim.x + threadIdx.x); ir <= nr; ir += (int)(gridDim.x * blockDim.x)) {
    if (ir != 0) {
      c[ir] = 1 / (dr * dr) * (1 + 1. / (2 * ir));
      d[ir] = 1 / (dr * dr) * (1 - 1. / (2 * ir));
    } else {
      c[ir] = d[ir] = 1 / (dr * dr);
    }
  }
}

}  // namespace

namespace spf {

// called by spf_ctx_create with the context's device current
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return SPF_EINVAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return SPF_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPF_EHIP;
  g_ptr[device] = (double *)p;
  return SPF_OK;
}

}  // namespace spf

extern "C" {

int spf_poisson_solve_noalloc(spf_ctx *c, double *Psi, const double *rhs, const int16_t *flags, int32_t nz,
                              int32_t nr, double dz, double dr, int32_t left_bc, int32_t right_bc, double accuracy,
                              int64_t iteration_limit, int64_t *iterations) {
  if (!c || !Psi || !flags || !iterations) return fail(c, SPF_EINVAL, "null argument");
  if (nz < 3 || nr < 3 || !(dz > 0) || !(dr > 0)) return fail(c, SPF_EINVAL, "nz, nr must be >= 3 and dz, dr > 0");
  if (accuracy <= 0) return fail(c, SPF_EINVAL, "accuracy is <= 0 in solve_poisson_cyl");
  if (iteration_limit <= 1) return fail(c, SPF_EINVAL, "iteration_limit is <= 1 in solve_poisson_cyl");
  double *S = c->device < kMaxDevices ? g_ptr[c->device] : nullptr;
  if (!S) return fail(c, SPF_EHIP, "the scratch array of this device was not resolved");
  // a run view of the caller's buffers and the array's regions: the SOR
  // launchers read only these fields
  Run R;
  R.d.nz = nz;
  R.d.nr = nr;
  R.d.dz = dz;
  R.d.dr = dr;
  R.d.left_bc = left_bc;
  R.d.right_bc = right_bc;
  R.ldz = (int)(((int64_t)nz + 1 + 31) / 32 * 32);
  R.cells = (int64_t)R.ldz * (nr + 1);
  const int64_t tiles = sor_tiles(R);
  if (R.cells > kMaxPoints || tiles > kMaxTiles || nr + 1 > kRows || sor_mask_words(R) > kMaxTiles * kTileWords)
    return fail(c, SPF_EINVAL, "grid larger than the no-allocation scratch array holds (ldz*(nr+1) <= " +
                                   std::to_string(kMaxPoints) + ", SOR tiles <= " + std::to_string(kMaxTiles) + ")");
  if (iteration_limit > kMaxIter)
    return fail(c, SPF_EINVAL, "iteration_limit above the no-allocation limit " + std::to_string(kMaxIter));
  if (sor_batch(R) > kPartialRows) return fail(c, SPF_EINVAL, "SOR batch deeper than the partial rows");
  R.flags = const_cast<int16_t *>(flags);
  R.sor_masks = reinterpret_cast<uint32_t *>(S + OFF_MASKS);
  R.red = S + OFF_PARTIALS;
  R.red_n = kPartialRows * kMaxTiles;
  R.omega = S + OFF_OMEGA;
  R.sor_c = S + OFF_C;
  R.sor_d = S + OFF_D;
  R.sor_sums = S + OFF_SUMS;
  R.sor_state = reinterpret_cast<int64_t *>(S + OFF_STATE);
  R.red_counter = reinterpret_cast<int32_t *>(S + OFF_COUNTER);
  R.stage = S + OFF_DUMP;
  R.sor_c_host0 = 1 / (dr * dr);
  hipStream_t s = c->stream;
  HIPCALL(c, hipSetDevice(c->device));
  // tables: omega only when the array does not hold this schedule already
  const double PI_D = 3.141592653589793238462643;
  const double r_jacobi = (cos(PI_D / nz) + (dz / dr) * (dz / dr) * cos(PI_D / nr)) / (1 + (dz / dr) * (dz / dr));
  {
    std::lock_guard<std::mutex> lock(g_mu);
    const bool have = g_omega_n[c->device] >= iteration_limit && g_omega_rj[c->device] == r_jacobi;
    hipLaunchKernelGGL(sor_tables_kernel, dim3(1), dim3(256), 0, s, r_jacobi, have ? (int64_t)0 : iteration_limit,
                       R.omega, nr, dr, R.sor_c, R.sor_d);
    HIPCALL(c, hipGetLastError());
    if (!have) {
      g_omega_rj[c->device] = r_jacobi;
      g_omega_n[c->device] = iteration_limit;
    }
  }
  HIPCALL(c, launch_sor_reset(R, s));
  HIPCALL(c, launch_sor_masks(R, s));
  if (rhs) HIPCALL(c, launch_sum_abs(rhs, R.ldz, nz, nr - 1, R.red, R.red_n, R.sor_sums, s));
  // Batches ping-pong between the caller's Psi and the array's spare buffer,
  // each followed by its stopping-rule check; once a check stops the solve,
  // later batches return without writing, so the stopping batch's input is
  // still intact and it is re-run for exactly the half-sweeps needed.
  const int K = sor_batch(R);
  double *buf[2] = {Psi, S + OFF_SPARE};
  int cur = 0;                 // buf[cur] holds Psi before half-sweep `start`
  int64_t start = 0;
  bool drive_serial = false;
  int out = 0;
  *iterations = -1;
  int64_t *flag = c->pinned + 56;
  for (;;) {
    int64_t converged = -1, ambiguous = 0, n0 = start, b = 0;
    while (n0 < iteration_limit && converged < 0) {
      const int k = (int)std::min<int64_t>(K, iteration_limit - n0);
      HIPCALL(c, launch_sor_batch(R, buf[(cur + b) & 1], buf[(cur + b + 1) & 1], rhs, n0, k, 0, 0, s));
      HIPCALL(c, launch_sor_check(R, 0, n0, k, accuracy, s));
      n0 += k;
      b++;
      if (b % 16 == 0 || n0 >= iteration_limit) {
        HIPCALL(c, hipMemcpyAsync(flag, R.sor_state, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCALL(c, hipStreamSynchronize(s));
        if (flag[0]) {
          converged = flag[1];
          ambiguous = flag[2];
        }
      }
    }
    if (converged < 0) {
      out = (cur + (int)(b & 1)) & 1;
      break;
    }
    const int64_t bc = (converged - start) / K, nc0 = start + bc * K;
    const double *in = buf[(cur + bc) & 1];
    out = (cur + (int)bc + 1) & 1;
    if (!ambiguous) {
      HIPCALL(c, launch_sor_batch(R, in, buf[out], rhs, nc0, (int)(converged - nc0 + 1), 0, 1, s));
      *iterations = converged;
      break;
    }
    // within rounding of the threshold: Psi before the half-sweep, the
    // half-sweep written as the reference writes it, and the serial sums
    if (converged > nc0)
      HIPCALL(c, launch_sor_batch(R, in, buf[out], rhs, nc0, (int)(converged - nc0), 0, 1, s));
    else
      HIPCALL(c, hipMemcpyAsync(buf[out], in, sizeof(double) * (size_t)R.cells, hipMemcpyDeviceToDevice, s));
    HIPCALL(c, launch_sor_exact_sweep(R, buf[out], rhs, converged, R.stage, s));
    HIPCALL(c, launch_serial_sum_ir_iz(R.stage, nz, nz, nr - 1, R.sor_sums + 8, s));
    if (rhs && !drive_serial) HIPCALL(c, launch_serial_sum_iz_ir(rhs, R.ldz, nz, nr - 1, 1, R.sor_sums + 9, s));
    drive_serial = rhs != nullptr;
    HIPCALL(c, hipMemcpyAsync(flag, R.sor_sums + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
    double res, drive = 0;
    memcpy(&res, flag, sizeof(double));
    if (rhs) memcpy(&drive, flag + 1, sizeof(double));
    if (res < accuracy * drive || res == 0) {   // src/poisson.c:125-126
      *iterations = converged;
      break;
    }
    if (converged + 1 >= iteration_limit) break;
    cur = out;
    start = converged + 1;
    HIPCALL(c, launch_sor_state_clear(R, s));
  }
  if (buf[out] != Psi)
    HIPCALL(c, hipMemcpyAsync(Psi, buf[out], sizeof(double) * (size_t)R.cells, hipMemcpyDeviceToDevice, s));
  HIPCALL(c, hipStreamSynchronize(s));
  return SPF_OK;
}

}  // extern "C"
