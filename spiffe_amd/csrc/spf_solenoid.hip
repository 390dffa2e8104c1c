// spf_solenoid.hip -- imposed solenoid fields (process_solenoid_definition,
// src/solenoid.c:18-127, and B_solenoid / B_ring, src/Bsolenoid.c:17-86).
//
// Every grid point sums the fields of `turns` current rings, each an exact
// Landau-Lifshitz ring field with complete elliptic integrals: O(points x
// turns) work with a dozen square roots per ring, one thread per point.  The
// points' coordinates are the reference's running sums z += dz, r += dr, made
// on the host.  An extension adds a Gaussian on-axis profile with its
// first-order off-axis expansion.
#include <cfloat>
#include <cmath>
#include <vector>

#include "spf_internal.h"

using namespace spf;

namespace {

constexpr double PI_D = 3.141592653589793238462643;

// Complete elliptic integrals K(k) and E(k) of modulus k by the arithmetic-
// geometric mean (Abramowitz & Stegun 17.6.1-17.6.4): a0 = 1, b0 = sqrt(1-k^2),
// c0 = k; K = pi/(2 a_N), E = K (1 - sum 2^(n-1) c_n^2).  The loop stops when
// |a - b| <= 2^-52 a or after 40 steps, the same rule on host and device.
__host__ __device__ inline void ellip_ke(double k, double *K, double *E) {
  double a = 1.0, b = sqrt(1.0 - k * k), c = k;
  double sum = 0.5 * c * c, p = 0.5;
  for (int n = 0; n < 40 && fabs(a - b) > 2.220446049250313e-16 * a; n++) {
    c = (a - b) / 2;
    double t = (a + b) / 2;
    b = sqrt(a * b);
    a = t;
    p *= 2;
    sum += p * c * c;
  }
  *K = PI_D / (2 * a);
  *E = *K * (1 - sum);
}

// B_ring, src/Bsolenoid.c:55-86: field of 1 A in a ring of radius `radius`
__host__ __device__ inline void b_ring(double *B_rho, double *B_z, double rho, double z, double radius) {
  double radius2 = radius * radius;
  double z2 = z * z;
  double r2 = rho * rho;
  double s1 = radius2 + r2 + z2;
  double s2 = radius2 - r2 - z2;
  radius2 = radius + rho;
  double s3 = radius2 * radius2 + z2;
  radius2 = radius - rho;
  double s4 = radius2 * radius2 + z2;
  double k = sqrt(4 * radius * rho / s3);
  double K, E;
  ellip_ke(k, &K, &E);
  double F = 2e-7 / sqrt(s3);   // mu_o/(2 pi)
  if (rho)
    *B_rho = F * z / rho * (-K + E * s1 / s4);
  else
    *B_rho = 0;
  *B_z = F * (K + E * s2 / s4);
}

// B_solenoid, src/Bsolenoid.c:17-48: `coils` equispaced rings, mirrored
// about z = 0 when symmetry is +1 or -1
__host__ __device__ inline void b_solenoid(double *B_rho, double *B_z, double rho, double z, double radius,
                                           double zStart, double zEnd, int64_t coils, int64_t symmetry) {
  double dzCoil = coils > 1 ? (zEnd - zStart) / (coils - 1) : 0;
  double Brho1, Bz1;
  *B_rho = *B_z = 0;
  for (int64_t i = 0; i < coils; i++) {
    double zCoil = zStart + i * dzCoil;
    b_ring(&Brho1, &Bz1, rho, z - zCoil, radius);
    *B_rho += Brho1;
    *B_z += Bz1;
    if (symmetry && zCoil) {
      b_ring(&Brho1, &Bz1, rho, z + zCoil, radius);
      if (symmetry == 1) {
        *B_rho += Brho1;
        *B_z += Bz1;
      } else if (symmetry == -1) {
        *B_rho -= Brho1;
        *B_z -= Bz1;
      }
    }
  }
}

struct SolArgs {
  int nz, ldz, nr_limit;
  const double *z, *r;      // running-sum coordinates, nz+1 and nr_limit+1 of them
  double radius, z_start, z_end, current;
  int64_t turns, symmetry;
  double *Bz, *Br;
};

__global__ void __launch_bounds__(256) solenoid_kernel(SolArgs a) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(a.nz + 1) * (a.nr_limit + 1);
  if (t >= n) return;
  int iz = (int)(t % (a.nz + 1)), ir = (int)(t / (a.nz + 1));
  double Br1, Bz1;
  b_solenoid(&Br1, &Bz1, a.r[ir], a.z[iz], a.radius, a.z_start, a.z_end, a.turns, a.symmetry);
  int64_t o = fix(a.ldz, iz, ir);
  a.Br[o] += Br1 * a.current;
  a.Bz[o] += Bz1 * a.current;
}

// Gaussian extension: Bz += G(z), Br += -(r/2) G'(z), with G and G' per row
__global__ void __launch_bounds__(256) gaussian_kernel(int nz, int ldz, int nr_limit, const double *G,
                                                       const double *Gp, const double *r, double *Bz, double *Br) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(nz + 1) * (nr_limit + 1);
  if (t >= n) return;
  int iz = (int)(t % (nz + 1)), ir = (int)(t / (nz + 1));
  int64_t o = fix(ldz, iz, ir);
  Bz[o] += G[iz];
  Br[o] += -(r[ir] / 2) * Gp[iz];
}

__global__ void __launch_bounds__(256) scale_rows_kernel(int nz, int ldz, int nr_limit, double f, double *Bz,
                                                         double *Br) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)(nz + 1) * (nr_limit + 1);
  if (t >= n) return;
  int64_t o = fix(ldz, (int)(t % (nz + 1)), (int)(t / (nz + 1)));
  Bz[o] *= f;
  Br[o] *= f;
}

}  // namespace

extern "C" {

void spf_solenoid_defaults(spf_solenoid_desc *d) {
  if (!d) return;
  *d = spf_solenoid_desc{};
  d->turns = 1;   // src/solenoid.c:30-33
}

int spf_solenoid_add(spf_ctx *c, int32_t run, const spf_solenoid_desc *d, double *current_used) {
  Run *R = get_run(c, run);
  if (!R || !d) return fail(c, SPF_EINVAL, "bad run id or null solenoid description");
  // checks of src/solenoid.c:37-47
  if (!d->gaussian && d->radius <= 0) return fail(c, SPF_EINVAL, "radius <= 0 is not valid");
  if (d->evaluation_radius_limit < 0) return fail(c, SPF_EINVAL, "evaluation_radius_limit < 0  is not valid");
  if (d->z_start > d->z_end) return fail(c, SPF_EINVAL, "z_start > z_end is not valid");
  if (d->turns <= 0) return fail(c, SPF_EINVAL, "turns <= 0 is not valid");
  if (d->symmetry && d->symmetry != -1 && d->symmetry != 1) return fail(c, SPF_EINVAL, "symmetry must be 0, 1, or -1");
  if (d->gaussian && !(d->gaussian_sigma > 0)) return fail(c, SPF_EINVAL, "gaussian_sigma must be positive");
  if (d->turns > (1 << 20)) return fail(c, SPF_EINVAL, "turns too large");
  const int nz = R->d.nz, nr = R->d.nr;
  const double dz = R->d.dz, dr = R->d.dr;
  int64_t nrLimit;
  if (!d->evaluation_radius_limit)
    nrLimit = nr;
  else if ((nrLimit = (int64_t)(d->evaluation_radius_limit / dr + 0.5)) > nr)
    nrLimit = nr;
  // running sums z += dz, r += dr, src/solenoid.c:73-83
  std::vector<double> zs(nz + 1), rs(nrLimit + 1);
  double z = R->d.zmin, r = 0;
  for (int iz = 0; iz <= nz; iz++, z += dz) zs[iz] = z;
  for (int64_t ir = 0; ir <= nrLimit; ir++, r += dr) rs[ir] = r;
  double *dzs = nullptr, *drs = nullptr, *dG = nullptr;
  HIPCALL(c, hipMalloc(&dzs, sizeof(double) * (nz + 1)));
  HIPCALL(c, hipMalloc(&drs, sizeof(double) * (nrLimit + 1)));
  HIPCALL(c, hipMemcpyAsync(dzs, zs.data(), sizeof(double) * (nz + 1), hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipMemcpyAsync(drs, rs.data(), sizeof(double) * (nrLimit + 1), hipMemcpyHostToDevice, c->stream));
  double *Bz = R->BzImp, *Br = R->BrImp;
  const int64_t npts = (int64_t)(nz + 1) * (nrLimit + 1);
  const unsigned blocks = (unsigned)((npts + 255) / 256);
  double current = d->current;
  int rc = SPF_OK;
  double *h = (double *)(c->pinned + 32);
  if (!d->gaussian) {
    if (d->bucking) {
      // the bucking current cancels the field already on the axis at z_buck
      int64_t iz = (int64_t)((d->z_buck - R->d.zmin) / dz);
      if (iz < 0 || iz > nz) {
        rc = fail(c, SPF_EINVAL, "z_buck outside the grid");
      } else {
        double Br1, Bz1;
        b_solenoid(&Br1, &Bz1, 0.0, d->z_buck, d->radius, d->z_start, d->z_end, d->turns, d->symmetry);
        HIPCALL(c, hipMemcpyAsync(h, Bz + fix(R->ldz, (int)iz, 0), sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(c, hipStreamSynchronize(c->stream));
        current = Bz1 != 0 ? -h[0] / Bz1 : 0;
      }
    }
    if (rc == SPF_OK) {
      SolArgs a{nz, R->ldz, (int)nrLimit, dzs, drs, d->radius, d->z_start, d->z_end, current, d->turns,
                d->symmetry, Bz, Br};
      hipLaunchKernelGGL(solenoid_kernel, dim3(blocks), dim3(256), 0, c->stream, a);
      HIPCALL(c, hipGetLastError());
    }
  } else {
    // Gaussian on-axis profile of peak Bz_peak (an extension: the reference's
    // solenoid is coil-only); G and G' come from libm on the host
    std::vector<double> G(2 * (nz + 1));
    const double s2 = d->gaussian_sigma * d->gaussian_sigma;
    for (int iz = 0; iz <= nz; iz++) {
      double u = zs[iz] - d->gaussian_center;
      double g = d->Bz_peak * exp(-u * u / (2 * s2));
      G[iz] = g;
      G[nz + 1 + iz] = -u / s2 * g;
    }
    HIPCALL(c, hipMalloc(&dG, sizeof(double) * G.size()));
    HIPCALL(c, hipMemcpyAsync(dG, G.data(), sizeof(double) * G.size(), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(gaussian_kernel, dim3(blocks), dim3(256), 0, c->stream, nz, R->ldz, (int)nrLimit, dG,
                       dG + nz + 1, drs, Bz, Br);
    HIPCALL(c, hipGetLastError());
  }
  if (rc == SPF_OK && d->Bz_peak && !d->gaussian) {
    // scale to Bz_peak at the on-axis maximum, src/solenoid.c:84-110
    std::vector<double> axis(nz + 1);
    HIPCALL(c, hipMemcpyAsync(axis.data(), Bz, sizeof(double) * (nz + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    double Bz_max = -DBL_MAX;
    int64_t iz_max = -1;
    for (int iz = 0; iz <= nz; iz++)
      if (axis[iz] > Bz_max) {
        iz_max = iz;
        Bz_max = axis[iz];
      }
    if (iz_max == -1)
      rc = fail(c, SPF_EINVAL, "problem scaling solenoid fields to Bz_peak---didn't find maximum!");
    else if (Bz_max == 0)
      rc = fail(c, SPF_EINVAL, "problem scaling solenoid fields to Bz_peak---maximum field is zero");
    else {
      hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks), dim3(256), 0, c->stream, nz, R->ldz, (int)nrLimit,
                         d->Bz_peak / Bz_max, Bz, Br);
      HIPCALL(c, hipGetLastError());
    }
  }
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipFree(dzs));
  HIPCALL(c, hipFree(drs));
  if (dG) HIPCALL(c, hipFree(dG));
  if (rc) return rc;
  R->d.has_solenoid = 1;   // the gather adds BzImposed/BrImposed from now on
  if (current_used) *current_used = current;
  return SPF_OK;
}

}  // extern "C"
