// spf_poisson.hip -- Poisson correction (src/poisson_correction.c:73-454),
// the Chebyshev red-black SOR solver (src/poisson.c:17-130) and the imposed
// electrostatic field (src/poisson_correction.c:456-562).
//
// Red-black SOR: within one half-sweep every updated point reads only points
// of the other colour, so the serial loop order of the reference does not
// affect Psi; the device updates one colour per launch and gets the same
// bits.  The convergence test needs sum |residual| after every half-sweep:
// each block writes a partial sum, the last block to finish adds the partials
// in block order (fixed, so deterministic) and sets a done flag that makes
// the remaining queued half-sweeps return at once.  The host checks the flag
// once per batch of launches.  The sum's rounding differs from the serial
// sum, so the iteration count can differ from the reference only when the
// residual lands within rounding of the threshold.
#include "spf_internal.h"

namespace spf {

namespace {

constexpr int SB = 256;  // SOR block: 64 (iz pairs) x 4 (ir)

struct SorArgs {
  int nz, nr, ldz, left_dir, right_dir;
  double a, e, accuracy;
  const double *c, *d;   // per-row stencil coefficients
  const double *omega;   // per-half-sweep omega
  const double *rhs;     // may be null
  double *Psi;
  const int16_t *flags;
  double *partials;
  int32_t *counter;
  int64_t *state;        // [0] done, [1] iterations
  double *sums;          // [0] drive sum, [1] residual sum of the last half-sweep
  int64_t n;
  int tiles_k, nblocks;
};

__global__ __launch_bounds__(SB) void sor_kernel(SorArgs a) {
  __shared__ double sh[SB];
  __shared__ int last;
  if (a.state[0]) return;
  const int color = (int)(a.n & 1);
  int tk = blockIdx.x % a.tiles_k, tr = blockIdx.x / a.tiles_k;
  int k = tk * 64 + (threadIdx.x & 63);
  int ir = tr * 4 + (threadIdx.x >> 6);
  double ares = 0;
  if (ir < a.nr - 1) {
    int iz = 2 * k + ((color + ir) & 1);   // (iz + ir) % 2 == n % 2
    bool in_range = (iz >= 1 && iz <= a.nz - 2) || (iz == 0 && a.left_dir && a.nz - 1 > 0) ||
                    (iz == a.nz - 1 && a.right_dir && iz > 0);
    if (iz < a.nz && in_range) {
      int64_t p = fix(a.ldz, iz, ir);
      if (!(a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) {
        const double *P = a.Psi;
        double up = P[fix(a.ldz, iz, ir + 1)];
        double zs;
        if (iz == 0)
          zs = P[fix(a.ldz, 1, ir)] + P[fix(a.ldz, 1, ir)];
        else if (iz == a.nz - 1)
          zs = P[fix(a.ldz, iz - 1, ir)] + P[fix(a.ldz, iz - 1, ir)];
        else
          zs = P[fix(a.ldz, iz + 1, ir)] + P[fix(a.ldz, iz - 1, ir)];
        double drive = a.rhs ? a.rhs[p] : 0.0;
        double c = a.c[ir], residual, omega = a.omega[a.n];
        if (ir == 0) {
          residual = a.a * zs + 4 * c * up + (a.e - 2 * c) * P[p] - drive;
          a.Psi[p] = P[p] - omega * residual / (a.e - 2 * c);
        } else {
          double dn = P[fix(a.ldz, iz, ir - 1)];
          residual = a.a * zs + c * up + a.d[ir] * dn + a.e * P[p] - drive;
          a.Psi[p] = P[p] - omega * residual / a.e;
        }
        ares = fabs(residual);
      }
    }
  }
  // fixed-order block sum
  sh[threadIdx.x] = ares;
  __syncthreads();
  for (int off = SB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.partials[blockIdx.x] = sh[0];
    __threadfence();
    int t = atomicAdd(a.counter, 1);
    last = (t == a.nblocks - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  volatile const double *vp = a.partials;
  double s = 0;
  for (int i = threadIdx.x; i < a.nblocks; i += SB) s += vp[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = SB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double total = sh[0];
    *a.counter = 0;
    a.sums[1] = total;
    if (a.n > 0 && (total < a.accuracy * a.sums[0] || total == 0)) {
      a.state[0] = 1;
      a.state[1] = a.n;
    }
  }
}

struct QArgs {
  int nz, nr, ldz, left_neu, right_neu, upper_neu;
  double eps0;
  const double *Ez, *Er;
  const double *Ar_in, *Ar_out, *Az, *volume;
  const int16_t *flags;
  double *Q;
};

// Gauss-law apparent charge per cell, src/poisson_correction.c:141-182
__global__ __launch_bounds__(256) void gauss_q_kernel(QArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)a.nz * a.nr;
  if (i >= n) return;
  int ir = (int)(i / a.nz), iz = (int)(i % a.nz);
  int64_t p = fix(a.ldz, iz, ir);
  double dQ;
  double ez = (a.Ez[fix(a.ldz, iz + 1, ir)] - a.Ez[p]) * a.Az[ir];
  if (ir == 0)
    dQ = (a.Er[fix(a.ldz, iz, ir + 1)] * a.Ar_out[ir] + ez) * a.eps0;
  else
    dQ = (a.Er[fix(a.ldz, iz, ir + 1)] * a.Ar_out[ir] - a.Er[p] * a.Ar_in[ir] + ez) * a.eps0;
  bool surface = (a.flags[p] & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE)) || (iz == 0 && a.left_neu) ||
                 (iz == a.nz - 1 && a.right_neu) || (ir == a.nr - 1 && a.upper_neu);
  a.Q[p] = surface ? 0.0 : dQ;
}

// Q -> -rho/eps0 driving term, :267-274
__global__ __launch_bounds__(256) void q_rhs_kernel(QArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)a.nz * a.nr;
  if (i >= n) return;
  int ir = (int)(i / a.nz), iz = (int)(i % a.nz);
  int64_t p = fix(a.ldz, iz, ir);
  a.Q[p] = -a.Q[p] / (a.volume[ir] * a.eps0);
}

__global__ __launch_bounds__(256) void zero_psi_kernel(double *Psi, int nz, int nr, int ldz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)nz * nr;
  if (i >= n) return;
  int ir = (int)(i / nz), iz = (int)(i % nz);
  Psi[fix(ldz, iz, ir)] = 0.0;
}

struct EArgs {
  int nz, nr, ldz, left_neu, right_neu, imposed;
  double dz, dr, ff;
  const double *Psi;
  const double *Ez, *Er;  // old (correction) or current imposed arrays
  double *Ezn, *Ern;      // new (correction: other ping-pong buffer; imposed: same arrays)
  const int16_t *flags;
};

// correction: Ez -= -(dPsi)/dz (src/poisson_correction.c:290-308);
// imposed: Ez = -ff*(dPsi)/dz (:524-543); masked points keep their value
__device__ inline double ez_upd(const EArgs &a, int iz, int ir) {
  int64_t p = fix(a.ldz, iz, ir);
  if (a.flags[p] & SPF_FL_EZ_ZERO) return a.Ez[p];
  double d = a.Psi[p] - a.Psi[fix(a.ldz, iz - 1, ir)];
  if (a.imposed) return -a.ff * d / a.dz;
  return a.Ez[p] - -d / a.dz;
}

__device__ inline double er_upd(const EArgs &a, int iz, int ir) {
  int64_t p = fix(a.ldz, iz, ir);
  if (a.flags[p] & SPF_FL_ER_ZERO) return a.Er[p];
  double d = a.Psi[p] - a.Psi[fix(a.ldz, iz, ir - 1)];
  if (a.imposed) return -a.ff * d / a.dr;
  return a.Er[p] - -d / a.dr;
}

__global__ __launch_bounds__(256) void e_from_psi_kernel(EArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(a.nz + 1) * (a.nr + 1);
  if (i >= n) return;
  int ir = (int)(i / (a.nz + 1)), iz = (int)(i % (a.nz + 1));
  int64_t p = fix(a.ldz, iz, ir);
  // Ez: loop ir < nr, iz 1..nz-1; ends mirrored; row nr-1 zeroed
  double ez;
  if (ir == a.nr - 1)
    ez = 0.0;
  else if (ir == a.nr)
    ez = a.Ez[p];
  else if (iz == 0) {
    double v = ez_upd(a, 1, ir);
    ez = a.left_neu ? v : -v;
  } else if (iz == a.nz) {
    double v = ez_upd(a, a.nz - 1, ir);
    ez = a.right_neu ? v : -v;
  } else
    ez = ez_upd(a, iz, ir);
  // Er: loop ir 1..nr, iz < nz; axis mirror; Neumann ends zeroed
  double er;
  if (iz == a.nz)
    er = a.Er[p];
  else if (a.left_neu && iz == 0)
    er = 0.0;
  else if (a.right_neu && iz == a.nz - 1)
    er = 0.0;
  else if (ir == 0)
    er = -er_upd(a, iz, 1);
  else
    er = er_upd(a, iz, ir);
  a.Ezn[p] = ez;
  a.Ern[p] = er;
}

__global__ void sor_reset_kernel(int64_t *state, int32_t *counter, double *sums) {
  state[0] = 0;
  state[1] = -1;
  *counter = 0;
  sums[0] = 0;
  sums[1] = 0;
}

QArgs q_args(Run &R) {
  QArgs q;
  q.nz = R.d.nz;
  q.nr = R.d.nr;
  q.ldz = R.ldz;
  q.left_neu = R.d.left_bc == SPF_NEUMANN;
  q.right_neu = R.d.right_bc == SPF_NEUMANN;
  q.upper_neu = R.d.upper_bc == SPF_NEUMANN;
  q.eps0 = R.d.epsilon_o;
  q.Ez = R.E2[0][R.cur[0]];
  q.Er = R.E2[1][R.cur[1]];
  q.Ar_in = R.volume + (R.d.nr + 1);
  q.Ar_out = R.volume + 2 * (R.d.nr + 1);
  q.Az = R.volume + 3 * (R.d.nr + 1);
  q.volume = R.volume;
  q.flags = R.flags;
  q.Q = R.Q;
  return q;
}

}  // namespace

hipError_t launch_sor_reset(Run &R, hipStream_t s) {
  hipLaunchKernelGGL(sor_reset_kernel, dim3(1), dim3(1), 0, s, R.sor_state, R.red_counter, R.sor_sums);
  return hipGetLastError();
}

int64_t sor_blocks(const Run &R) {
  int tiles_k = ((R.d.nz + 1) / 2 + 63) / 64;
  int tiles_r = (R.d.nr - 1 + 3) / 4;
  return (int64_t)tiles_k * tiles_r;
}

hipError_t launch_sor_halfsweep(Run &R, double *Psi, const double *rhs, int64_t n, double accuracy, hipStream_t s) {
  SorArgs a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_dir = R.d.left_bc == SPF_DIRICHLET;
  a.right_dir = R.d.right_bc == SPF_DIRICHLET;
  double dz = R.d.dz, dr = R.d.dr;
  a.a = 1 / (dz * dz);
  a.e = -2 * (1 / (dz * dz) + 1 / (dr * dr));
  a.accuracy = accuracy;
  a.c = R.sor_c;
  a.d = R.sor_d;
  a.omega = R.omega;
  a.rhs = rhs;
  a.Psi = Psi;
  a.flags = R.flags;
  a.partials = R.red;
  a.counter = R.red_counter;
  a.state = R.sor_state;
  a.sums = R.sor_sums;
  a.n = n;
  a.tiles_k = ((R.d.nz + 1) / 2 + 63) / 64;
  a.nblocks = (int)sor_blocks(R);
  hipLaunchKernelGGL(sor_kernel, dim3((unsigned)a.nblocks), dim3(SB), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_gauss_q(Run &R, hipStream_t s) {
  QArgs q = q_args(R);
  int64_t n = (int64_t)q.nz * q.nr;
  hipLaunchKernelGGL(gauss_q_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_q_to_rhs(Run &R, hipStream_t s) {
  QArgs q = q_args(R);
  int64_t n = (int64_t)q.nz * q.nr;
  hipLaunchKernelGGL(q_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q);
  return hipGetLastError();
}

hipError_t launch_zero_psi(Run &R, hipStream_t s) {
  int64_t n = (int64_t)R.d.nz * R.d.nr;
  hipLaunchKernelGGL(zero_psi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R.Psi, R.d.nz, R.d.nr,
                     R.ldz);
  return hipGetLastError();
}

hipError_t launch_field_from_psi(Run &R, int imposed, double field_factor, hipStream_t s) {
  EArgs a;
  a.nz = R.d.nz;
  a.nr = R.d.nr;
  a.ldz = R.ldz;
  a.left_neu = R.d.left_bc == SPF_NEUMANN;
  a.right_neu = R.d.right_bc == SPF_NEUMANN;
  a.imposed = imposed;
  a.dz = R.d.dz;
  a.dr = R.d.dr;
  a.ff = field_factor;
  a.flags = R.flags;
  if (imposed) {
    a.Psi = R.PsiImp;
    a.Ez = R.EzImp;
    a.Er = R.ErImp;
    a.Ezn = R.EzImp;
    a.Ern = R.ErImp;
  } else {
    a.Psi = R.Psi;
    a.Ez = R.E2[0][R.cur[0]];
    a.Er = R.E2[1][R.cur[1]];
    a.Ezn = R.E2[0][1 - R.cur[0]];
    a.Ern = R.E2[1][1 - R.cur[1]];
  }
  int64_t n = (int64_t)(a.nz + 1) * (a.nr + 1);
  hipLaunchKernelGGL(e_from_psi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !imposed) {
    R.cur[0] = 1 - R.cur[0];
    R.cur[1] = 1 - R.cur[1];
  }
  return e;
}

}  // namespace spf
