/*
 * spf_oracle_solenoid.c -- CPU restatement of spiffe's imposed solenoid field
 * (process_solenoid_definition, src/solenoid.c:18-127; B_solenoid and B_ring,
 * src/Bsolenoid.c:17-86).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's solenoid entry point;
 * only tests/ load it.
 *
 * The SDDS toolkit's K_cei/E_cei are not available; the complete elliptic
 * integrals here are the arithmetic-geometric-mean method of Abramowitz &
 * Stegun 17.6 (stop when |a-b| <= 2^-52 a, at most 40 steps).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spf_oracle.h"

#define S_PI 3.141592653589793238462643

void oracle_ellip_ke(double k, double *K, double *E) {
  double a = 1.0, b = sqrt(1.0 - k * k), c = k, sum, pow2 = 0.5, an;
  int n;
  sum = 0.5 * c * c;
  for (n = 0; n < 40 && fabs(a - b) > 2.220446049250313e-16 * a; n++) {
    c = (a - b) / 2;
    an = (a + b) / 2;
    b = sqrt(a * b);
    a = an;
    pow2 *= 2;
    sum += pow2 * c * c;
  }
  *K = S_PI / (2 * a);
  *E = *K * (1 - sum);
}

/* src/Bsolenoid.c:55-86 (Landau & Lifshitz ring field, 1 A) */
void oracle_b_ring(double *B_rho, double *B_z, double rho, double z, double radius) {
  double radius2, z2, r2, K, E, F, s1, s2, s3, s4, k;
  radius2 = radius * radius;
  z2 = z * z;
  r2 = rho * rho;
  s1 = radius2 + r2 + z2;
  s2 = radius2 - r2 - z2;
  radius2 = radius + rho;
  s3 = radius2 * radius2 + z2;
  radius2 = radius - rho;
  s4 = radius2 * radius2 + z2;
  k = sqrt(4 * radius * rho / s3);
  oracle_ellip_ke(k, &K, &E);
  F = 2e-7 / sqrt(s3);
  if (rho)
    *B_rho = F * z / rho * (-K + E * s1 / s4);
  else
    *B_rho = 0;
  *B_z = F * (K + E * s2 / s4);
}

/* src/Bsolenoid.c:17-48 */
void oracle_b_solenoid(double *B_rho, double *B_z, double rho, double z, double radius, double zStart, double zEnd,
                       int64_t coils, int64_t symmetry) {
  int64_t iCoil;
  double dzCoil, Brho1, Bz1, zCoil;
  *B_rho = *B_z = 0;
  if (coils > 1)
    dzCoil = (zEnd - zStart) / (coils - 1);
  else
    dzCoil = 0;
  for (iCoil = 0; iCoil < coils; iCoil++) {
    zCoil = zStart + iCoil * dzCoil;
    oracle_b_ring(&Brho1, &Bz1, rho, z - zCoil, radius);
    *B_rho += Brho1;
    *B_z += Bz1;
    if (symmetry && zCoil) {
      oracle_b_ring(&Brho1, &Bz1, rho, z + zCoil, radius);
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

static int sbomb(char *err, int64_t errlen, const char *msg) {
  if (err && errlen > 0)
    snprintf(err, (size_t)errlen, "%s", msg);
  return 1;
}

#define AT(a, iz, ir) (a)[(iz) * (f->nr + 1) + (ir)]

/* process_solenoid_definition, src/solenoid.c:18-127, adding into
 * f->BzImposed / f->BrImposed; the Gaussian on-axis profile with
 * Br = -(r/2) dBz/dz is the library's extension, restated here */
int oracle_define_solenoid(ofields *f, const osolenoid_desc *d, double *current_used, char *err, int64_t errlen) {
  int64_t iz, ir, nz = f->nz, nr = f->nr, nrLimit;
  double z, r, dz = f->dz, dr = f->dr, Br1, Bz1, current = d->current;
  if (!f->BzImposed || !f->BrImposed)
    return sbomb(err, errlen, "BzImposed/BrImposed arrays are missing");
  if (!d->gaussian && d->radius <= 0)
    return sbomb(err, errlen, "radius <= 0 is not valid");
  if (d->evaluation_radius_limit < 0)
    return sbomb(err, errlen, "evaluation_radius_limit < 0  is not valid");
  if (d->z_start > d->z_end)
    return sbomb(err, errlen, "z_start > z_end is not valid");
  if (d->turns <= 0)
    return sbomb(err, errlen, "turns <= 0 is not valid");
  if (d->symmetry && d->symmetry != -1 && d->symmetry != 1)
    return sbomb(err, errlen, "symmetry must be 0, 1, or -1");
  if (!d->evaluation_radius_limit)
    nrLimit = nr;
  else if ((nrLimit = (int64_t)(d->evaluation_radius_limit / dr + 0.5)) > nr)
    nrLimit = nr;
  if (!d->gaussian) {
    if (d->bucking) {
      iz = (int64_t)((d->z_buck - f->zmin) / dz);
      oracle_b_solenoid(&Br1, &Bz1, 0.0, d->z_buck, d->radius, d->z_start, d->z_end, d->turns, d->symmetry);
      current = Bz1 != 0 ? -AT(f->BzImposed, iz, 0) / Bz1 : 0;
    }
    z = f->zmin;
    for (iz = 0; iz <= nz; iz++) {
      r = 0;
      for (ir = 0; ir <= nrLimit; ir++) {
        oracle_b_solenoid(&Br1, &Bz1, r, z, d->radius, d->z_start, d->z_end, d->turns, d->symmetry);
        AT(f->BrImposed, iz, ir) += Br1 * current;
        AT(f->BzImposed, iz, ir) += Bz1 * current;
        r += dr;
      }
      z += dz;
    }
    if (d->Bz_peak) {
      double Bz_max = -DBL_MAX;
      int64_t iz_max = -1;
      for (iz = 0; iz <= nz; iz++)
        if (AT(f->BzImposed, iz, 0) > Bz_max) {
          iz_max = iz;
          Bz_max = AT(f->BzImposed, iz, 0);
        }
      if (iz_max == -1 || Bz_max == 0)
        return sbomb(err, errlen, "problem scaling solenoid fields to Bz_peak");
      for (iz = 0; iz <= nz; iz++)
        for (ir = 0; ir <= nrLimit; ir++) {
          AT(f->BzImposed, iz, ir) *= d->Bz_peak / Bz_max;
          AT(f->BrImposed, iz, ir) *= d->Bz_peak / Bz_max;
        }
    }
  } else {
    double s2 = d->gaussian_sigma * d->gaussian_sigma;
    z = f->zmin;
    for (iz = 0; iz <= nz; iz++) {
      double u = z - d->gaussian_center;
      double g = d->Bz_peak * exp(-u * u / (2 * s2));
      double gp = -u / s2 * g;
      r = 0;
      for (ir = 0; ir <= nrLimit; ir++) {
        AT(f->BzImposed, iz, ir) += g;
        AT(f->BrImposed, iz, ir) += -(r / 2) * gp;
        r += dr;
      }
      z += dz;
    }
  }
  if (current_used)
    *current_used = current;
  return 0;
}
