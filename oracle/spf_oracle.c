/*
 * spf_oracle.c -- CPU restatement of spiffe's per-time-step EM-PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it; the HIP product path never does.
 *
 * Each function restates the cited reference lines with the same
 * floating-point expression order, so that a correctly-rounded FP64 machine
 * with no FMA contraction produces the same bits.  Build: gcc -O2
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Deliberate differences from the reference (all recorded in the design notes):
 *  - the reference's compaction defect (np truncated to the first lost index,
 *    src/advance_particles.c:491-518) is reproduced when compat != 0; with
 *    compat == 0 the swap-remove also moves the lost flag;
 *  - the screen-interpolation quirk that uses the LAST particle's old momenta
 *    (src/advance_particles.c:79, :202-204, :380-385) is reproduced when
 *    compat != 0 and replaced by per-particle old momenta otherwise;
 *  - screen buffers are allocated per call for every screen (the reference
 *    allocates only screen 0 once, :344-360), never reproduced;
 *  - test-beam forces, on-axis field expansions, secondaries and SDDS output
 *    are not part of this path; on-axis fields contribute an explicit +0.0.
 */
#include "spf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793238462643
#define IDX(f, iz, ir) ((int64_t)(iz) * ((int64_t)(f)->nr + 1) + (int64_t)(ir))
#define A(f, arr, iz, ir) ((arr)[IDX(f, iz, ir)])

static inline double sqr(double x) { return x * x; }
static inline int64_t isqr(int64_t i) { return i * i; }

/* Flag derivation from the metal drawing, src/geometry.c:319-362. */
void oracle_calculate_flags(int16_t *mf, int32_t nz, int32_t nr) {
  int64_t iz, ir, w = (int64_t)nr + 1;
#define MF(a, b) mf[(int64_t)(a) * w + (b)]
  for (iz = 0; iz < nz + 1; iz++) {
    for (ir = 0; ir < nr + 1; ir++) {
      if (!(MF(iz, ir) & OFL_IS_METAL))
        continue;
      MF(iz, ir) |= OFL_EPHI_ZERO;
      if (iz && MF(iz - 1, ir) & OFL_IS_METAL) {
        MF(iz, ir) |= OFL_EZ_ZERO;
        MF(iz, ir) |= OFL_BR_ZERO;
      }
      if (ir == 0 || (ir && MF(iz, ir - 1) & OFL_IS_METAL)) {
        MF(iz, ir) |= OFL_ER_ZERO;
        MF(iz, ir) |= OFL_BZ_ZERO;
      }
      if (iz && MF(iz - 1, ir) & OFL_IS_METAL &&
          (ir == 0 || (MF(iz, ir - 1) & OFL_IS_METAL && MF(iz - 1, ir - 1) & OFL_IS_METAL)))
        MF(iz, ir) |= OFL_BPHI_ZERO;
    }
  }
  for (ir = 0; ir < nr; ir++) {
    iz = nz - 1;
    while (iz >= 0 && MF(iz, ir) & OFL_IS_INSIDE)
      iz--;
    if (iz != 0 && (iz += 1) < nz - 1)
      MF(iz, ir) |= OFL_LAST_ACTIVE_PT;
  }
#undef MF
}

/* src/integrate.c:210-241 */
void oracle_zero_current_arrays(ofields *f) {
  int64_t iz, ir;
  if (f->modes & OFL_TM_FIELDS)
    for (iz = 0; iz <= f->nz; iz++)
      for (ir = 0; ir <= f->nr; ir++)
        A(f, f->Jz, iz, ir) = A(f, f->Jr, iz, ir) = 0;
  if (f->modes & OFL_TE_FIELDS)
    for (iz = 0; iz <= f->nz; iz++)
      for (ir = 0; ir <= f->nr; ir++)
        A(f, f->Jphi, iz, ir) = 0;
}

/* src/integrate.c:243-275 */
void oracle_multiply_current_arrays(ofields *f, double factor) {
  int64_t iz, ir;
  if (f->modes & OFL_TM_FIELDS)
    for (iz = 0; iz <= f->nz; iz++)
      for (ir = 0; ir <= f->nr; ir++) {
        A(f, f->Jz, iz, ir) *= factor;
        A(f, f->Jr, iz, ir) *= factor;
      }
  if (f->modes & OFL_TE_FIELDS)
    for (iz = 0; iz <= f->nz; iz++)
      for (ir = 0; ir <= f->nr; ir++)
        A(f, f->Jphi, iz, ir) *= factor;
}

/* FDTD leapfrog, src/integrate.c:277-541 */
void oracle_advance_field_solutions(ofields *f) {
  int64_t iz, ir, nz = f->nz, nr = f->nr;
  double dr = f->dr, dz = f->dz, dti = f->dti, dB;
  double eps0 = f->epsilon_o, mu0 = f->mu_o;
  double rp, rm, factor, cp, cm;
  const int16_t *mf = f->metal_flags;
#define FL(a, b) mf[IDX(f, a, b)]

  if (f->modes & OFL_TM_FIELDS) {
    double *Ez = f->Ez, *Er = f->Er, *Bphi = f->Bphi, *Bphi2 = f->Bphi2, *Jz = f->Jz, *Jr = f->Jr;
    /* -- Ez, :323-359 */
    factor = 1 / (eps0 * mu0);
#pragma omp parallel for private(iz, rp, rm, cp, cm) schedule(static)
    for (ir = 0; ir < nr; ir++) {
      rp = (ir + 0.5) * dr;
      rm = (ir - 0.5) * dr;
      if (ir != 0) {
        cp = 2 * factor / (sqr(rp) - sqr(rm));
        cm = rm * cp;
        cp = rp * cp;
      } else {
        cp = 4 * factor / dr;
        cm = 0;
      }
      for (iz = 1; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_EZ_ZERO))
          A(f, Ez, iz, ir) += dti * (-A(f, Jz, iz, ir) / eps0 + (cp * A(f, Bphi, iz, ir + 1) - cm * A(f, Bphi, iz, ir)));
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
      if (f->left_bc == O_NEUMANN)
        A(f, Ez, 0, ir) = A(f, Ez, 1, ir);
      else
        A(f, Ez, 0, ir) = -A(f, Ez, 1, ir);
      if (f->right_bc == O_NEUMANN)
        A(f, Ez, nz, ir) = A(f, Ez, nz - 1, ir);
      else
        A(f, Ez, nz, ir) = -A(f, Ez, nz - 1, ir);
    }
    for (iz = 0; iz <= nz; iz++)
      A(f, Ez, iz, nr - 1) = 0;

    /* -- Er, :361-380 */
    factor = 1 / (mu0 * eps0 * dz);
#pragma omp parallel for private(iz) schedule(static)
    for (ir = 1; ir < nr; ir++) {
      for (iz = 0; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_ER_ZERO))
          A(f, Er, iz, ir) += dti * (-A(f, Jr, iz, ir) / eps0 - factor * (A(f, Bphi, iz + 1, ir) - A(f, Bphi, iz, ir)));
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
    }
    for (iz = 0; iz < nz; iz++)
      A(f, Er, iz, 0) = -A(f, Er, iz, 1);
    if (f->left_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Er, 0, ir) = 0;
    if (f->right_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Er, nz - 1, ir) = 0;

    /* -- Bphi, :382-432 */
    factor = (f->rmax - f->dr / 2) / (f->rmax + f->dr / 2);
#pragma omp parallel for private(iz, dB) schedule(static)
    for (ir = 1; ir < nr; ir++) {
      for (iz = 1; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_BPHI_ZERO)) {
          dB = dti * ((A(f, Ez, iz, ir) - A(f, Ez, iz, ir - 1)) / dr - (A(f, Er, iz, ir) - A(f, Er, iz - 1, ir)) / dz);
          A(f, Bphi2, iz, ir) = A(f, Bphi, iz, ir) + dB / 2;
          A(f, Bphi, iz, ir) += dB;
        }
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
    }
    for (iz = 1; iz < nz; iz++) {
      A(f, Bphi, iz, 0) = -A(f, Bphi, iz, 1);
      A(f, Bphi2, iz, 0) = -A(f, Bphi2, iz, 1);
      A(f, Bphi, iz, nr - 1) = A(f, Bphi, iz, nr - 2) * factor;
      A(f, Bphi2, iz, nr - 1) = A(f, Bphi2, iz, nr - 2) * factor;
    }
    if (f->left_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++) {
        A(f, Bphi, 0, ir) = A(f, Bphi, 1, ir);
        A(f, Bphi2, 0, ir) = A(f, Bphi2, 1, ir);
      }
    else
      for (ir = 0; ir <= nr; ir++) {
        A(f, Bphi, 0, ir) = -A(f, Bphi, 1, ir);
        A(f, Bphi2, 0, ir) = -A(f, Bphi2, 1, ir);
      }
    if (f->right_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++) {
        A(f, Bphi, nz, ir) = A(f, Bphi, nz - 1, ir);
        A(f, Bphi2, nz, ir) = A(f, Bphi2, nz - 1, ir);
      }
    else
      for (ir = 0; ir <= nr; ir++) {
        A(f, Bphi, nz, ir) = -A(f, Bphi, nz - 1, ir);
        A(f, Bphi2, nz, ir) = -A(f, Bphi2, nz - 1, ir);
      }
  }
  if (f->modes & OFL_TE_FIELDS) {
    double *Ephi = f->Ephi, *Jphi = f->Jphi, *Br = f->Br, *Br2 = f->Br2, *Bz = f->Bz, *Bz2 = f->Bz2;
    /* -- Ephi, :449-471 */
    factor = 1. / (eps0 * mu0);
#pragma omp parallel for private(iz) schedule(static)
    for (ir = 1; ir < nr; ir++)
      for (iz = 0; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_EPHI_ZERO))
          A(f, Ephi, iz, ir) += dti * (-A(f, Jphi, iz, ir) / eps0 +
                                       factor * ((A(f, Br, iz + 1, ir) - A(f, Br, iz, ir)) / dz - (A(f, Bz, iz, ir + 1) - A(f, Bz, iz, ir)) / dr));
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
    for (iz = 0; iz <= nz; iz++)
      A(f, Ephi, iz, 0) = 0;
    if (f->right_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Ephi, nz - 1, ir) = 0;
    if (f->left_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Ephi, 0, ir) = 0;
    if (f->upper_bc == O_NEUMANN)
      for (iz = 0; iz <= nz; iz++)
        A(f, Ephi, iz, nr - 1) = 0;

    /* -- Br, :473-509 */
#pragma omp parallel for private(iz, dB) schedule(static)
    for (ir = 0; ir < nr; ir++)
      for (iz = 1; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_BR_ZERO)) {
          dB = dti * (A(f, Ephi, iz, ir) - A(f, Ephi, iz - 1, ir)) / dz;
          A(f, Br2, iz, ir) = A(f, Br, iz, ir) + dB / 2;
          A(f, Br, iz, ir) += dB;
        }
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
    if (f->right_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++) {
        A(f, Br, nz, ir) = A(f, Br, nz - 1, ir);
        A(f, Br2, nz, ir) = A(f, Br2, nz - 1, ir);
      }
    else
      for (ir = 0; ir <= nr; ir++) {
        A(f, Br, nz, ir) = -A(f, Br, nz - 1, ir);
        A(f, Br2, nz, ir) = -A(f, Br2, nz - 1, ir);
      }
    if (f->left_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++) {
        A(f, Br, 0, ir) = A(f, Br, 1, ir);
        A(f, Br2, 0, ir) = A(f, Br2, 1, ir);
      }
    else
      for (ir = 0; ir <= nr; ir++) {
        A(f, Br, 0, ir) = -A(f, Br, 1, ir);
        A(f, Br2, 0, ir) = -A(f, Br2, 1, ir);
      }

    /* -- Bz, :511-539 */
#pragma omp parallel for private(iz, dB, factor) schedule(static)
    for (ir = 1; ir < nr; ir++) {
      factor = -2 * dti / (dr * (isqr(ir + 1) - isqr(ir)));
      for (iz = 0; iz < nz; iz++) {
        if (!(FL(iz, ir) & OFL_BZ_ZERO)) {
          dB = factor * ((ir + 1) * A(f, Ephi, iz, ir) - ir * A(f, Ephi, iz, ir - 1));
          A(f, Bz2, iz, ir) = A(f, Bz, iz, ir) + dB / 2;
          A(f, Bz, iz, ir) += dB;
        }
        if (FL(iz, ir) & OFL_LAST_ACTIVE_PT)
          break;
      }
    }
    for (iz = 0; iz <= nz; iz++) {
      A(f, Bz, iz, 0) = A(f, Bz, iz, 1);
      A(f, Bz2, iz, 0) = A(f, Bz2, iz, 1);
    }
    if (f->right_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Bz, nz - 1, ir) = A(f, Bz2, nz - 1, ir) = 0;
    if (f->left_bc == O_NEUMANN)
      for (ir = 0; ir <= nr; ir++)
        A(f, Bz, 0, ir) = A(f, Bz2, 0, ir) = 0;
  }
#undef FL
}

/* Line-current add of add_antenna_excitation, src/antenna.c:137-154 */
void oracle_add_line_current(ofields *f, int32_t direction, int64_t iposition,
                             int64_t istart, int64_t iend, double J) {
  int64_t i;
  if (direction == 0) {
    for (i = istart; i <= iend; i++)
      A(f, f->Jz, i, iposition) += J;
  } else if (direction == 1) {
    for (i = istart; i <= iend; i++)
      A(f, f->Jr, iposition, i) += J;
  } else if (direction == 2) {
    for (i = istart; i <= iend; i++)
      A(f, f->Jphi, i, iposition) += J;
  }
}

/* Waveform-enveloped antenna current, src/antenna.c:118-136.
 * INTERPOLATE(y1, y0, x1, x0, x) from the SDDS toolkit is linear
 * interpolation y0 + (y1-y0)/(x1-x0)*(x-x0); its exact expression order is
 * not pinned by the reference (the macro lives in the un-vendored mdb.h). */
double oracle_antenna_current(const double *t, const double *W, int64_t n_pts, int64_t *last_index,
                              double current, double frequency, double phase, double time,
                              int32_t *error) {
  double J;
  int64_t i;
  *error = 0;
  if (n_pts <= 1) {
    *error = 1;
    return 0;
  }
  if (time <= t[0])
    J = W[0];
  else if (time > t[n_pts - 1] || *last_index >= (n_pts - 1))
    J = W[n_pts - 1];
  else {
    if ((i = *last_index) != 0 && time < t[i]) {
      *error = 2;
      return 0;
    }
    while (i < n_pts - 2 && time >= t[i + 1])
      i++;
    *last_index = i;
    J = W[i] + (W[i + 1] - W[i]) / (t[i + 1] - t[i]) * (time - t[i]);
  }
  J *= current * sin(2 * PI_D * frequency * time + phase);
  return J;
}

/* Current deposition, src/space_charge.c:34-295.  With longitudinal smearing
 * a particle whose iz is nz adds fp*Jm and fp*Jp to row nz+1, one past the
 * reference's arrays (zero terms for Jz and Jr, whose fp is 0 there; for
 * Jphi a particle beyond zmax); those writes are left out here. */
void oracle_add_particle_currents(ofields *f, const obeam *b) {
  int64_t ip, iz, ir, nz = f->nz, nr = f->nr;
  double fp, fm, Jm, Jp, dz = f->dz, dr = f->dr;
  double Q_over_piDzDr2, zmin = f->zmin, one_over_piDzDr2, QVG;
  uint32_t modes = f->modes;
  one_over_piDzDr2 = 1 / (PI_D * dz * sqr(dr));

  if (modes & OFL_TM_FIELDS) {
    /* Jz, :61-139 */
    for (ip = 0; ip < b->np; ip++) {
      if (!(b->status[ip] & OPART_ACTIVE))
        continue;
      if (modes & OFL_RADIAL_SMEAR)
        ir = b->r[ip] / dr;
      else
        ir = b->r[ip] / dr + 0.5;
      if (modes & OFL_LONGIT_SMEAR)
        iz = (b->z[ip] - zmin) / dz + 0.5;
      else
        iz = (b->z[ip] - zmin) / dz + 1.0;
      if (iz < 0 || iz > nz)
        continue;
      if (ir < 0 || ir > nr)
        continue;
      Q_over_piDzDr2 = b->Q[ip] * one_over_piDzDr2;
      QVG = b->vz[ip] * Q_over_piDzDr2;
      if (modes & OFL_RADIAL_SMEAR)
        fm = 1 - (fp = b->r[ip] / dr - ir);
      else {
        fm = 1;
        fp = 0;
      }
      if (ir == 0) {
        Jm = 4 * fm * QVG;
        Jp = fp * QVG / (sqr(1.5) - sqr(0.5));
      } else if (ir < nr - 1) {
        Jm = fm * QVG / (sqr(ir + 0.5) - sqr(ir - 0.5));
        Jp = fp * QVG / (sqr(ir + 1.5) - sqr(ir + 0.5));
      } else {
        Jm = 0;
        Jp = fp * QVG / (sqr(ir + 0.5) - sqr(ir - 0.5));
        ir--;
      }
      if (modes & OFL_LONGIT_SMEAR) {
        if (iz == 0) {
          fp = 1;
          fm = 0;
        } else if (iz == nz) {
          fm = 1;
          fp = 0;
        } else
          fm = 1 - (fp = (b->z[ip] - zmin) / dz + 0.5 - iz);
      } else {
        fm = 1;
        fp = 0;
      }
      A(f, f->Jz, iz, ir) += fm * Jm;
      if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
        A(f, f->Jz, iz + 1, ir) += fp * Jm;
      if (modes & OFL_RADIAL_SMEAR) {
        A(f, f->Jz, iz, ir + 1) += fm * Jp;
        if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
          A(f, f->Jz, iz + 1, ir + 1) += fp * Jp;
      }
    }
    /* Jr, :141-215 */
    for (ip = 0; ip < b->np; ip++) {
      if (!(b->status[ip] & OPART_ACTIVE))
        continue;
      if (modes & OFL_RADIAL_SMEAR)
        ir = b->r[ip] / dr + 0.5;
      else
        ir = b->r[ip] / dr + 1.0;
      if (modes & OFL_LONGIT_SMEAR)
        iz = (b->z[ip] - zmin) / dz;
      else
        iz = (b->z[ip] - zmin) / dz + 0.5;
      if (iz < 0 || iz > nz)
        continue;
      if (ir < 0 || ir >= nr)
        continue;
      Q_over_piDzDr2 = b->Q[ip] * one_over_piDzDr2;
      if (modes & OFL_RADIAL_SMEAR)
        fm = 1 - (fp = b->r[ip] / dr + 0.5 - ir);
      else {
        fm = 1;
        fp = 0;
      }
      QVG = b->vr[ip] * Q_over_piDzDr2;
      if (ir == 0) {
        Jm = 0;
        Jp = QVG;
      } else if (ir < nr - 1) {
        Jm = fm * QVG / (sqr(ir) - sqr(ir - 1));
        Jp = fp * QVG / (sqr(ir + 1) - sqr(ir));
      } else {
        Jm = 0;
        ir--;
        Jp = QVG / (sqr(ir + 1) - sqr(ir));
      }
      if (modes & OFL_LONGIT_SMEAR) {
        if (iz == nz) {
          fm = 1;
          fp = 0;
        } else
          fm = 1 - (fp = (b->z[ip] - zmin) / dz - iz);
      } else {
        fm = 1;
        fp = 0;
      }
      A(f, f->Jr, iz, ir) += fm * Jm;
      if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
        A(f, f->Jr, iz + 1, ir) += fp * Jm;
      if (modes & OFL_RADIAL_SMEAR) {
        A(f, f->Jr, iz, ir + 1) += fm * Jp;
        if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
          A(f, f->Jr, iz + 1, ir + 1) += fp * Jp;
      }
    }
  }
  if (modes & OFL_TE_FIELDS) {
    /* Jphi, :222-293 */
    for (ip = 0; ip < b->np; ip++) {
      if (!(b->status[ip] & OPART_ACTIVE))
        continue;
      if (modes & OFL_RADIAL_SMEAR)
        ir = b->r[ip] / dr;
      else
        ir = b->r[ip] / dr + 0.5;
      if (modes & OFL_LONGIT_SMEAR)
        iz = (b->z[ip] - zmin) / dz;
      else
        iz = (b->z[ip] - zmin) / dz + 0.5;
      if (iz < 0 || iz > nz)
        continue;
      if (ir < 0 || ir > nr)
        continue;
      Q_over_piDzDr2 = b->Q[ip] * one_over_piDzDr2;
      if (modes & OFL_RADIAL_SMEAR)
        fm = 1 - (fp = b->r[ip] / dr - ir);
      else {
        fm = 1;
        fp = 0;
      }
      QVG = b->vphi[ip] * Q_over_piDzDr2;
      if (ir == 0) {
        Jm = 4 * fm * QVG;
        Jp = fp * QVG / (sqr(1.5) - sqr(0.5));
      } else if (ir < nr - 1) {
        Jm = fm * QVG / (sqr(ir + 0.5) - sqr(ir - 0.5));
        Jp = fp * QVG / (sqr(ir + 1.5) - sqr(ir + 0.5));
      } else {
        Jm = 0;
        Jp = fp * QVG / (sqr(ir + 0.5) - sqr(ir - 0.5));
        ir--;
      }
      if (modes & OFL_LONGIT_SMEAR) {
        if (iz == nz - 1)
          fm = 1 - (fp = 0);
        else
          fm = 1 - (fp = (b->z[ip] - zmin) / dz - iz);
      } else {
        fm = 1;
        fp = 0;
      }
      A(f, f->Jphi, iz, ir) += fm * Jm;
      if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
        A(f, f->Jphi, iz + 1, ir) += fp * Jm;
      if (modes & OFL_RADIAL_SMEAR) {
        A(f, f->Jphi, iz, ir + 1) += fm * Jp;
        if ((modes & OFL_LONGIT_SMEAR) && iz + 1 <= nz)
          A(f, f->Jphi, iz + 1, ir + 1) += fp * Jp;
      }
    }
  }
}

/* src/advance_particles.c:31-65 */
short oracle_check_loss_code(int64_t iz, int64_t ir, const ofields *f) {
  short lossCode = 0;
  int64_t iz1, ir1;
  const int16_t *mf = f->metal_flags;
  /* outside the flag array (inactive particle at r < -rmax): the reference
   * reads out of bounds here; both implementations count it as metal */
  if (iz < 0 || iz > f->nz || ir < 0 || ir > f->nr)
    return 3;
  if (mf[IDX(f, iz, ir)] & OFL_IS_INSIDE)
    lossCode = 3;
  else if (mf[IDX(f, iz, ir)] & OFL_IS_SURFACE) {
    iz1 = iz + 1;
    if (iz1 > f->nz)
      lossCode = 4;
    else {
      ir1 = ir + 1;
      if (ir1 > f->nr)
        lossCode = 5;
      else if (mf[IDX(f, iz1, ir1)] & OFL_IS_METAL && mf[IDX(f, iz, ir1)] & OFL_IS_METAL &&
               mf[IDX(f, iz1, ir)] & OFL_IS_METAL)
        lossCode = 6;
    }
  }
  return lossCode;
}

/* src/advance_particles.c:605-620 */
void oracle_set_particle_velocities(double *vz, double *vr, double *vphi, double *gamma,
                                    double pz, double pr, double r_pphi, double r, double c_mks) {
  if (r) {
    *gamma = sqrt(sqr(pz) + sqr(pr) + sqr(r_pphi / r) + 1);
    *vphi = c_mks * r_pphi / r / (*gamma);
  } else {
    *gamma = sqrt(sqr(pz) + sqr(pr) + 1);
    *vphi = 0;
  }
  *vz = c_mks * pz / (*gamma);
  *vr = c_mks * pr / (*gamma);
}

/* Staggered bilinear gather, src/advance_particles.c:622-890 */
void oracle_interpolate_fields(double *EzR, double *ErR, double *EphiR, double *BzR, double *BrR, double *BphiR,
                               double z, double r, const ofields *f) {
  int64_t iz, ir, nz = f->nz, nr = f->nr, iz1, ir1;
  int longit = (f->modes & OFL_LONGIT_INTERP) != 0, radial = (f->modes & OFL_RADIAL_INTERP) != 0;
  double dz = f->dz, dr = f->dr, fr, fl, fa, fb, zmin = f->zmin;
  double Ez, Er, Ephi, Bz, Br, Bphi;
  const int16_t *mf = f->metal_flags;
#define FL(a, b) mf[IDX(f, a, b)]

  Ez = Er = Bphi = Ephi = Br = Bz = 0;
  /* -- Ez and Br, :649-710 */
  if (longit) {
    iz = (fr = (z - zmin) / dz + 0.5);
    fr -= iz;
    fl = 1 - fr;
  } else {
    iz = (z - zmin) / dz + 1;
    fl = 1;
    fr = 0;
  }
  if (radial) {
    ir = (fa = r / dr);
    fa -= ir;
    fb = 1 - fa;
  } else {
    ir = r / dr + 0.5;
    fa = 0;
    fb = 1;
  }
  if (iz >= 0 && iz < nz && ir >= 0 && ir < nr) {
    if ((iz1 = iz + 1) >= nz)
      iz1 = iz;
    if ((ir1 = ir + 1) >= nr)
      ir1 = ir;
    if (FL(iz, ir + 1) & OFL_ER_ZERO && FL(iz, ir) & OFL_ER_ZERO) {
      if (iz > 0 && (FL(iz - 1, ir) || FL(iz - 1, ir + 1)))
        iz = iz1;
      else
        iz1 = iz;
    } else if (FL(iz1, ir + 1) & OFL_ER_ZERO && FL(iz1, ir) & OFL_ER_ZERO) {
      iz1 = iz;
    }
    if (f->modes & OFL_TM_FIELDS)
      Ez = fb * (fl * (A(f, f->Ez, iz, ir) + A(f, f->EzImposed, iz, ir)) +
                 fr * (A(f, f->Ez, iz1, ir) + A(f, f->EzImposed, iz1, ir))) +
           fa * (fl * (A(f, f->Ez, iz, ir1) + A(f, f->EzImposed, iz, ir1)) +
                 fr * (A(f, f->Ez, iz1, ir1) + A(f, f->EzImposed, iz1, ir1)));
    if (r != 0 && f->modes & OFL_TE_FIELDS)
      Br = fb * (fl * A(f, f->Br2, iz, ir) + fr * A(f, f->Br2, iz1, ir)) +
           fa * (fl * A(f, f->Br2, iz, ir1) + fr * A(f, f->Br2, iz1, ir1));
  }

  /* -- Er and Bz, :712-752 */
  if (longit) {
    iz = (fr = (z - zmin) / dz);
    fr -= iz;
    fl = 1 - fr;
  } else {
    iz = (z - zmin) / dz + 0.5;
    fr = 0;
    fl = 1;
  }
  if (radial) {
    ir = (fa = r / dr + 0.5);
    fa -= ir;
    fb = 1 - fa;
  } else {
    ir = r / dr + 1;
    fa = 0;
    fb = 1;
  }
  if (iz >= 0 && iz < nz && ir >= 0) {
    if (ir >= nr)
      ir = nr - 1;
    iz1 = iz + 1;
    ir1 = ir + 1;
    if (FL(iz + 1, ir) & OFL_EZ_ZERO && FL(iz, ir) & OFL_EZ_ZERO)
      ir = ir1;
    else if (FL(iz + 1, ir + 1) & OFL_EZ_ZERO && FL(iz, ir + 1) & OFL_EZ_ZERO)
      ir1 = ir;
    if (r != 0 && f->modes & OFL_TM_FIELDS)
      Er = fl * (fb * (A(f, f->Er, iz, ir) + A(f, f->ErImposed, iz, ir)) + fa * (A(f, f->Er, iz, ir1) + A(f, f->ErImposed, iz, ir1))) +
           fr * (fb * (A(f, f->Er, iz1, ir) + A(f, f->ErImposed, iz1, ir)) + fa * (A(f, f->Er, iz1, ir1) + A(f, f->ErImposed, iz1, ir1)));
    if (f->modes & OFL_TE_FIELDS)
      Bz = fl * (fb * A(f, f->Bz2, iz, ir) + fa * A(f, f->Bz2, iz, ir1)) + fr * (fb * A(f, f->Bz2, iz1, ir) + fa * A(f, f->Bz2, iz1, ir1));
  }

  /* -- Ephi, :754-786 */
  if (longit) {
    iz = (fr = (z - zmin) / dz);
    fr -= iz;
    fl = 1 - fr;
  } else {
    iz = (z - zmin) / dz + 0.5;
    fr = 0;
    fl = 1;
  }
  if (radial) {
    ir = (fa = r / dr);
    fa -= ir;
    fb = 1 - fa;
  } else {
    ir = r / dr + 0.5;
    fa = 0;
    fb = 1;
  }
  if (r != 0 && f->modes & OFL_TE_FIELDS && iz >= 0 && iz < nz && ir >= 0 && ir < nr) {
    if ((iz1 = iz + 1) >= nz)
      iz1 = iz;
    if ((ir1 = ir + 1) >= nr)
      ir1 = ir;
    Ephi = fl * fb * A(f, f->Ephi, iz, ir) + fr * fb * A(f, f->Ephi, iz1, ir) + fl * fa * A(f, f->Ephi, iz, ir1) + fr * fa * A(f, f->Ephi, iz1, ir1);
  }

  /* -- Bphi at time t, :788-825 */
  if (f->modes & OFL_TM_FIELDS) {
    if (longit) {
      iz = (fr = (z - zmin) / dz + 0.5);
      fr -= iz;
      fl = 1 - fr;
    } else {
      iz = (z - zmin) / dz + 1;
      fr = 0;
      fl = 1;
    }
    if (radial) {
      ir = (fa = r / dr + 0.5);
      fa -= ir;
      fb = 1 - fa;
    } else {
      ir = r / dr + 1;
      fa = 0;
      fb = 1;
    }
    if (r != 0 && iz >= 0 && iz < nz) {
      if ((iz1 = iz + 1) >= nz)
        iz1 = iz;
      if (ir >= nr)
        ir = nr - 1;
      if ((ir1 = ir + 1) >= nr)
        ir1 = ir;
      Bphi = fl * fb * A(f, f->Bphi2, iz, ir) + fr * fb * A(f, f->Bphi2, iz1, ir) + fa * fl * A(f, f->Bphi2, iz, ir1) + fr * fa * A(f, f->Bphi2, iz1, ir1);
    }
  }

  /* constants, :829-834; on-axis expansions contribute +0.0 when none are defined, :836-844 */
  Ez += f->constantEz;
  Er += f->constantEr;
  Ephi += f->constantEphi;
  Bz += f->constantBz;
  Br += f->constantBr;
  Bphi += f->constantBphi;
  Ez += 0.0;
  Er += 0.0;
  Bphi += 0.0;

  /* :846-857 */
  if (f->turnOffEz)
    Ez = 0;
  if (f->turnOffEr)
    Er = 0;
  if (f->turnOffEphi)
    Ephi = 0;
  if (f->turnOffBz)
    Bz = 0;
  if (f->turnOffBr)
    Br = 0;
  if (f->turnOffBphi)
    Bphi = 0;

  /* solenoid, :859-875 */
  if (f->BzImposed) {
    iz = (fr = (z - zmin) / dz);
    ir = (fa = r / dr);
    if (ir >= 0 && iz >= 0 && ir < nr && iz < nz) {
      fa -= ir;
      fb = 1 - fa;
      fr -= iz;
      fl = 1 - fr;
      Bz += fl * (fb * A(f, f->BzImposed, iz, ir) + fa * A(f, f->BzImposed, iz, ir + 1)) +
            fr * (fb * A(f, f->BzImposed, iz + 1, ir) + fa * A(f, f->BzImposed, iz + 1, ir + 1));
      Br += fl * (fb * A(f, f->BrImposed, iz, ir) + fa * A(f, f->BrImposed, iz, ir + 1)) +
            fr * (fb * A(f, f->BrImposed, iz + 1, ir) + fa * A(f, f->BrImposed, iz + 1, ir + 1));
    }
  }
  *EzR = Ez;
  *ErR = Er;
  *EphiR = Ephi;
  *BzR = Bz;
  *BrR = Br;
  *BphiR = Bphi;
#undef FL
}

/* src/secondaryEmission.c:239-372 */
/* flag lookup for the loss-coordinate search: the reference indexes with
 * iz1 = -1 when a particle ends below zmin (src/secondaryEmission.c:314) and
 * reads outside its array; both implementations read such points as 0 */
static int16_t loss_flag(const ofields *f, int64_t iz, int64_t ir) {
  if (iz < 0 || iz > f->nz || ir < 0 || ir > f->nr)
    return 0;
  return f->metal_flags[IDX(f, iz, ir)];
}

void oracle_find_loss_coordinates(double *zReturn, double *rReturn, short *directionCode,
                                  double z0, double r0, double z1, double r1, const ofields *f) {
  double zm, rm;
  int64_t iz1, ir1, iz0, ir0;
  *directionCode = 0;
  iz0 = (z0 - f->zmin) / f->dz;
  ir0 = r0 / f->dr;
  if (z1 < f->zmin)
    iz1 = -1;
  else
    iz1 = (z1 - f->zmin) / f->dz;
  if (r1 < 0)
    ir1 = -1;
  else
    ir1 = r1 / f->dr;
  if (r0 == r1 && z0 == z1) {
    *directionCode = 1;
    *zReturn = z0;
    *rReturn = r0;
    return;
  }
  if (ir0 == ir1) {
    *directionCode = iz0 > iz1 ? O_GOING_LEFT : O_GOING_RIGHT;
  } else if (iz0 == iz1) {
    *directionCode = ir0 > ir1 ? O_GOING_DOWN : O_GOING_UP;
  } else {
    if (z1 < z0) {
      if (z1 < 0)
        *directionCode = O_GOING_LEFT;
      else if (ir0 != ir1) {
        if (ir1 >= 0 && ir1 < f->nr && loss_flag(f, iz0, ir1) & OFL_IS_METAL)
          *directionCode = O_GOING_LEFT;
        else if (ir1 < 0 && loss_flag(f, iz0, -ir1) & OFL_IS_METAL)
          *directionCode = O_GOING_LEFT;
      } else if (ir0 == ir1 && (ir0 < (f->nr - 1) && loss_flag(f, iz0, ir0 + 1) & OFL_IS_METAL))
        *directionCode = O_GOING_LEFT;
      else
        *directionCode = r1 < r0 ? O_GOING_DOWN : O_GOING_UP;
    } else if (z1 > z0) {
      if (z1 > f->zmax)
        *directionCode = O_GOING_RIGHT;
      else if (ir0 != ir1) {
        if (ir1 >= 0 && ir1 < f->nr && loss_flag(f, iz1, ir1) & OFL_IS_METAL)
          *directionCode = O_GOING_RIGHT;
        else if (ir1 < 0 && loss_flag(f, iz1, -ir1) & OFL_IS_METAL)
          *directionCode = O_GOING_RIGHT;
      } else if (ir0 == ir1 && (ir0 < (f->nr - 1) && loss_flag(f, iz1, ir0 + 1) & OFL_IS_METAL))
        *directionCode = O_GOING_RIGHT;
      else
        *directionCode = r1 < r0 ? O_GOING_DOWN : O_GOING_UP;
    } else
      *directionCode = r1 < r0 ? O_GOING_DOWN : O_GOING_UP;
  }
  if (*directionCode == O_GOING_LEFT) {
    zm = ((int64_t)((z0 - f->zmin) / f->dz)) * f->dz + f->zmin + f->dz * 1e-12;
    rm = (r1 - r0) / (z1 - z0) * (zm - z0) + r0;
  } else if (*directionCode == O_GOING_RIGHT) {
    zm = ((int64_t)((z0 - f->zmin) / f->dz)) * f->dz + f->zmin + f->dz * (1 - 1e-12);
    rm = (r1 - r0) / (z1 - z0) * (zm - z0) + r0;
  } else if (*directionCode == O_GOING_DOWN) {
    rm = ((int64_t)(r0 / f->dr)) * f->dr + f->dr * 1e-12;
    zm = (z1 - z0) / (r1 - r0) * (rm - r0) + z0;
  } else {
    rm = ((int64_t)(r0 / f->dr)) * f->dr + f->dr * (1 - 1e-12);
    zm = (z1 - z0) / (r1 - r0) * (rm - r0) + z0;
  }
  *zReturn = zm;
  *rReturn = fabs(rm);
}

static void copy_particle(obeam *b, int64_t dst, int64_t src) {
  b->Q[dst] = b->Q[src];
  b->z[dst] = b->z[src];
  b->r[dst] = b->r[src];
  b->pz[dst] = b->pz[src];
  b->pr[dst] = b->pr[src];
  b->vz[dst] = b->vz[src];
  b->vr[dst] = b->vr[src];
  b->r_pphi[dst] = b->r_pphi[src];
  b->vphi[dst] = b->vphi[src];
  b->t0[dst] = b->t0[src];
  b->r0[dst] = b->r0[src];
  b->z0[dst] = b->z0[src];
  b->rmin[dst] = b->rmin[src];
  b->zmin[dst] = b->zmin[src];
  b->rmax[dst] = b->rmax[src];
  b->zmax[dst] = b->zmax[src];
  b->generation[dst] = b->generation[src];
  b->particleID[dst] = b->particleID[src];
  b->status[dst] = b->status[src];
}

/* Particle push with gather, loss, screens and compaction, src/advance_particles.c:67-575.
 * Returns 0, or 1 if a row buffer was too small (rows beyond the limit are dropped). */
int oracle_advance_particles(ofields *f, obeam *b, const oscreen *screens, int32_t n_screens,
                             double zForcesStart, int32_t compat,
                             oscreen_row *rows, int64_t max_rows, int64_t *n_rows,
                             olost_row *lost, int64_t max_lost, int64_t *n_lost) {
  int64_t ip, np = b->np, newActiveCount = 0;
  double dt = f->dti, dz = f->dz, dr = f->dr, zmin = f->zmin, zmax = f->zmax, rmax = f->rmax;
  double pz0 = 0, pr0 = 0, r_pphi0 = 0;
  double c_mks = f->c_mks;
  int overflow = 0;
  int64_t nalloc = np > 0 ? np : 1;
  char *is_lost = calloc(nalloc, 1), *r_was_negative = calloc(nalloc, 1);
  double *z0Save = malloc(sizeof(double) * nalloc), *r0Save = malloc(sizeof(double) * nalloc);
  double *z0 = malloc(sizeof(double) * nalloc), *vzz = malloc(sizeof(double) * nalloc);
  double *pphi_a = malloc(sizeof(double) * nalloc), *vr_a = malloc(sizeof(double) * nalloc);
  double *pz0a = malloc(sizeof(double) * nalloc), *pr0a = malloc(sizeof(double) * nalloc);
  double *rpphi0a = malloc(sizeof(double) * nalloc);
  *n_rows = 0;
  *n_lost = 0;

  /* main loop, :131-339; particles are independent, and in a threaded build
   * pz0/pr0/r_pphi0 end with the last particle's values as in the serial loop */
#pragma omp parallel for reduction(+ : newActiveCount) lastprivate(pz0, pr0, r_pphi0) schedule(static)
  for (ip = 0; ip < np; ip++) {
    double z, r, vz, vr, pz, pr, r_pphi, vphi, gamma, pphi, adjustFactor = 1;
    double Ez, Er, Ephi, Bz, Br, Bphi;
    int64_t iz, ir;
    short lossCode;
    z = z0[ip] = z0Save[ip] = b->z[ip];
    r = r0Save[ip] = b->r[ip];
    if (z > b->zmax[ip])
      b->zmax[ip] = z;
    if (z < b->zmin[ip])
      b->zmin[ip] = z;
    if (r > b->rmax[ip])
      b->rmax[ip] = r;
    if (r < b->rmin[ip])
      b->rmin[ip] = r;
    vz = b->vz[ip];
    vr = b->vr[ip];
    z += vz * dt;
    r += vr * dt;
    if (!(b->status[ip] & OPART_ACTIVE)) {
      iz = (z - zmin) / dz;
      ir = fabs(r) / dr;
      if (!(!(z >= f->zmin && z <= f->zmax && r <= f->rmax) || oracle_check_loss_code(iz, ir, f))) {
        b->status[ip] |= OPART_ACTIVE;
        newActiveCount++;
        b->t0[ip] = f->time;
        b->r0[ip] = b->r[ip];
        b->z0[ip] = b->z[ip];
        if (iz > 0 && oracle_check_loss_code(iz - 1, ir, f))
          adjustFactor = (z - (iz * f->dz + f->zmin)) / (z - z0[ip]);
        else
          adjustFactor = (z - f->zmin) / (z - z0[ip]);
        if (adjustFactor > 1)
          adjustFactor = 1;
      }
    }
    if (zForcesStart > f->zmin && zForcesStart < f->zmax) {
      if (z0[ip] < zForcesStart) {
        if (z < zForcesStart)
          adjustFactor = 0;
        else
          adjustFactor = (z - zForcesStart) / (z - z0[ip]);
      }
    }
    pz = pz0 = b->pz[ip];
    pr = pr0 = b->pr[ip];
    r_pphi = r_pphi0 = b->r_pphi[ip];
    pz0a[ip] = pz0;
    pr0a[ip] = pr0;
    rpphi0a[ip] = r_pphi0;
    vphi = b->vphi[ip];

    is_lost[ip] = 0;
    lossCode = 0;
    if ((b->status[ip] & OPART_ACTIVE) && (z < zmin || z > zmax || r > rmax)) {
      lossCode = 1;
      is_lost[ip] = 1;
    }
    iz = (z - zmin) / dz;
    ir = fabs(r) / dr;
    if ((b->status[ip] & OPART_ACTIVE) && (iz > f->nz || ir > f->nr || iz < 0 || ir < 0)) {
      lossCode = 2;
      is_lost[ip] = 1;
    }
    if (!is_lost[ip] && (b->status[ip] & OPART_ACTIVE)) {
      if ((lossCode = oracle_check_loss_code(iz, ir, f)))
        is_lost[ip] = 1;
    }
    (void)lossCode;
    r_was_negative[ip] = 0;
    if (r < 0) {
      r *= -1;
      pr *= -1;
      r_was_negative[ip] = 1;
    }
    if (!is_lost[ip] && (b->status[ip] & OPART_ACTIVE)) {
      if (z > b->zmax[ip])
        b->zmax[ip] = z;
      if (z < b->zmin[ip])
        b->zmin[ip] = z;
      if (r > b->rmax[ip])
        b->rmax[ip] = r;
      if (r < b->rmin[ip])
        b->rmin[ip] = r;
    }
    if (b->status[ip] & OPART_ACTIVE)
      oracle_interpolate_fields(&Ez, &Er, &Ephi, &Bz, &Br, &Bphi, z, r, f);
    else
      Ez = Er = Ephi = Bz = Br = Bphi = 0;
    /* momentum kick, :282-287 */
    pphi = (r ? r_pphi / r : 0);
    gamma = sqrt(sqr(pr) + sqr(pz) + sqr(pphi) + 1);
    pr += adjustFactor * dt * ((r ? sqr(pphi) * c_mks / gamma / r : 0) + b->QoMC * (Er + vphi * Bz - vz * Bphi)) / b->stiffness;
    pz += adjustFactor * dt * b->QoMC * (Ez + vr * Bphi - vphi * Br) / b->stiffness;
    r_pphi += adjustFactor * dt * b->QoMC * r * (Ephi + vz * Br - vr * Bz) / b->stiffness;

    b->z[ip] = z;
    b->r[ip] = r;
    b->pz[ip] = pz;
    b->pr[ip] = pr;
    b->r_pphi[ip] = r_pphi;
    oracle_set_particle_velocities(&b->vz[ip], &b->vr[ip], &b->vphi[ip], &gamma, pz, pr, r_pphi, r, c_mks);
    if (is_lost[ip]) {
      b->status[ip] &= ~OPART_ACTIVE;
      b->status[ip] |= OPART_LOST;
    }
    vzz[ip] = vz;
    pphi_a[ip] = pphi;
    vr_a[ip] = vr;
  }
  b->npActive += newActiveCount;

  /* screens, :362-434 (rows grouped per screen, in particle order within a screen) */
  {
    int32_t is;
    for (is = 0; is < n_screens; is++) {
      for (ip = 0; ip < np; ip++) {
        double dtp, r0, pphi0, spz0, spr0, srp0;
        int32_t js;
        /* r0 persists across the screen loop of one particle (:366-367, :408) */
        r0 = r0Save[ip];
        for (js = 0; js <= is; js++) {
          int fwd = screens[js].z_position <= b->z[ip] && screens[js].z_position > z0[ip] && screens[js].direction == 1;
          int bwd = !fwd && screens[js].z_position >= b->z[ip] && screens[js].z_position < z0[ip] && screens[js].direction == -1;
          oscreen_row row;
          if (!fwd && !bwd)
            continue;
          dtp = (screens[js].z_position - z0[ip]) / vzz[ip];
          if (fwd) {
            if (js != is)
              continue;
            spz0 = compat ? pz0 : pz0a[ip];
            spr0 = compat ? pr0 : pr0a[ip];
            srp0 = compat ? r_pphi0 : rpphi0a[ip];
            pphi_a[ip] = b->r[ip] ? b->r_pphi[ip] / b->r[ip] : 0;
            pphi0 = r0 ? srp0 / r0 : 0;
            row.r = r0 + vr_a[ip] * dtp;
            row.pz = spz0 + (b->pz[ip] - spz0) / dt * dtp;
            row.pr = spr0 + (b->pr[ip] - spr0) / dt * dtp;
            row.pphi = pphi0 + (pphi_a[ip] - pphi0) / dt * dtp;
          } else {
            r0 += vr_a[ip] * dtp;
            if (js != is)
              continue;
            row.r = r0;
            row.pz = b->pz[ip];
            row.pr = b->pr[ip];
            row.pphi = r0 ? b->r_pphi[ip] / r0 : 0;
          }
          row.t = (f->time - dt + dtp);
          row.q = b->Q[ip];
          row.r0 = b->r0[ip];
          row.t0 = b->t0[ip];
          row.rmin = b->rmin[ip];
          row.rmax = b->rmax[ip];
          row.zmin = b->zmin[ip];
          row.zmax = b->zmax[ip];
          row.generation = b->generation[ip];
          row.particleID = b->particleID[ip];
          row.ip = ip;
          row.screen = (int16_t)is;
          if (*n_rows < max_rows)
            rows[*n_rows] = row;
          else
            overflow = 1;
          (*n_rows)++;
        }
      }
    }
  }

  /* lost-particle rows, :436-488 */
  for (ip = 0; ip < np; ip++) {
    if (is_lost[ip]) {
      double zLost, rLost;
      short directionCode;
      olost_row row;
      oracle_find_loss_coordinates(&zLost, &rLost, &directionCode, z0Save[ip], r0Save[ip], b->z[ip],
                                   b->r[ip] * (r_was_negative[ip] ? -1 : 1), f);
      row.t = f->time;
      row.rLost = (float)rLost;
      row.zLost = (float)zLost;
      row.pz = (float)b->pz[ip];
      row.pr = (float)b->pr[ip];
      row.r0 = (float)b->r0[ip];
      row.z0 = (float)b->z0[ip];
      row.q = (float)b->Q[ip];
      row.generation = b->generation[ip];
      row.particleID = b->particleID[ip];
      row.ip = ip;
      row.directionCode = directionCode;
      row.pad = 0;
      if (*n_lost < max_lost)
        lost[*n_lost] = row;
      else
        overflow = 1;
      (*n_lost)++;
    }
  }

  /* compaction, :490-518 */
  for (ip = 0; ip < b->np; ip++) {
    if (is_lost[ip]) {
      b->np -= 1;
      b->npActive -= 1;
      copy_particle(b, ip, b->np);
      if (!compat)
        is_lost[ip] = is_lost[b->np];
      ip--;
    }
  }

  free(is_lost);
  free(r_was_negative);
  free(z0Save);
  free(r0Save);
  free(z0);
  free(vzz);
  free(pphi_a);
  free(vr_a);
  free(pz0a);
  free(pr0a);
  free(rpphi0a);
  return overflow;
}

/* Chebyshev-accelerated red-black SOR, src/poisson.c:17-130.
 * Psi/driving_term/flags use the (nz+1) x (nr+1) layout. */
int64_t oracle_solve_poisson_cyl(double *Psi, const double *driving_term, const int16_t *flags,
                                 int32_t left_bc, int32_t right_bc, int64_t nz, int64_t nr,
                                 double dz, double dr, double accuracy, int64_t iteration_limit,
                                 double *abs_total_residual_out) {
  int64_t iz, ir, n, w = nr + 1;
  double residual, omega, r_jacobi;
  double abs_total_driving_term, abs_total_residual = 0;
  double a, c = 0, d = 0, e;
#define P(a_, b_) Psi[(a_) * w + (b_)]
#define D(a_, b_) (driving_term ? driving_term[(a_) * w + (b_)] : 0)
#define F(a_, b_) flags[(a_) * w + (b_)]
  if (accuracy <= 0 || iteration_limit <= 1)
    return -3;
  abs_total_driving_term = 0.0;
  if (driving_term)
    for (iz = 0; iz < nz; iz++)
      for (ir = 0; ir < nr - 1; ir++)
        abs_total_driving_term += fabs(driving_term[iz * w + ir]);
  r_jacobi = (cos(PI_D / nz) + sqr(dz / dr) * cos(PI_D / nr)) / (1 + sqr(dz / dr));
  omega = 1.0;
  a = 1 / sqr(dz);
  e = -2 * (1 / sqr(dz) + 1 / sqr(dr));
  for (n = 0; n < iteration_limit; n++) {
    abs_total_residual = 0;
    /* a threaded build splits the rows of a half-sweep (each point reads only
     * the other colour) and adds the residual per thread: Psi is the same,
     * the residual total can round differently */
#pragma omp parallel for private(iz, c, d, residual) reduction(+ : abs_total_residual) schedule(static)
    for (ir = 0; ir < nr - 1; ir++) {
      if (ir != 0) {
        c = 1 / sqr(dr) * (1 + 1. / (2 * ir));
        d = 1 / sqr(dr) * (1 - 1. / (2 * ir));
      } else
        c = d = 1 / sqr(dr);
      if (left_bc == O_DIRICHLET) {
        iz = 0;
        if ((ir + iz) % 2 == n % 2 && !(F(iz, ir) & OFL_IS_METAL)) {
          if (ir == 0) {
            residual = a * (P(iz + 1, ir) + P(iz + 1, ir)) + 4 * c * P(iz, ir + 1) + (e - 2 * c) * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / (e - 2 * c);
          } else {
            residual = a * (P(iz + 1, ir) + P(iz + 1, ir)) + c * P(iz, ir + 1) + d * P(iz, ir - 1) + e * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / e;
          }
          abs_total_residual += fabs(residual);
        }
      }
      for (iz = 1; iz < nz - 1; iz++) {
        if ((ir + iz) % 2 == n % 2 && !(F(iz, ir) & OFL_IS_METAL)) {
          if (ir == 0) {
            residual = a * (P(iz + 1, ir) + P(iz - 1, ir)) + 4 * c * P(iz, ir + 1) + (e - 2 * c) * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / (e - 2 * c);
          } else {
            residual = a * (P(iz + 1, ir) + P(iz - 1, ir)) + c * P(iz, ir + 1) + d * P(iz, ir - 1) + e * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / e;
          }
          abs_total_residual += fabs(residual);
        }
      }
      if (right_bc == O_DIRICHLET) {
        iz = nz - 1;
        if ((ir + iz) % 2 == n % 2 && !(F(iz, ir) & OFL_IS_METAL)) {
          if (ir == 0) {
            residual = a * (P(iz - 1, ir) + P(iz - 1, ir)) + 4 * c * P(iz, ir + 1) + (e - 2 * c) * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / (e - 2 * c);
          } else {
            residual = a * (P(iz - 1, ir) + P(iz - 1, ir)) + c * P(iz, ir + 1) + d * P(iz, ir - 1) + e * P(iz, ir) - D(iz, ir);
            P(iz, ir) -= omega * residual / e;
          }
          abs_total_residual += fabs(residual);
        }
      }
    }
    if (n == 0)
      omega = 1.0 / (1 - sqr(r_jacobi) / 2);
    else
      omega = 1.0 / (1 - sqr(r_jacobi) * omega / 4);
    if (n > 0 && (abs_total_residual < accuracy * abs_total_driving_term || abs_total_residual == 0)) {
      if (abs_total_residual_out)
        *abs_total_residual_out = abs_total_residual;
      return n;
    }
  }
  if (abs_total_residual_out)
    *abs_total_residual_out = abs_total_residual;
  return -1;
#undef P
#undef D
#undef F
}

/* The omega used by half-sweep k (k = 0..n-1), same recurrence as src/poisson.c:47-49, :121-124. */
void oracle_poisson_omega_schedule(int64_t nz, int64_t nr, double dz, double dr, int64_t n, double *omega) {
  double r_jacobi = (cos(PI_D / nz) + sqr(dz / dr) * cos(PI_D / nr)) / (1 + sqr(dz / dr));
  double w = 1.0;
  int64_t k;
  for (k = 0; k < n; k++) {
    omega[k] = w;
    if (k == 0)
      w = 1.0 / (1 - sqr(r_jacobi) / 2);
    else
      w = 1.0 / (1 - sqr(r_jacobi) * w / 4);
  }
}

static void gauss_law_charge(ofields *f, double *volume, double *Qsurface, double *centroid, int first_pass) {
  int64_t iz, ir, nz = f->nz, nr = f->nr;
  double dr = f->dr, dz = f->dz, dr2 = sqr(dr), drdz = dr * dz, eps = f->epsilon_o;
  double Az, Ar_inner, Ar_outer, dQ;
  for (ir = 0; ir < nr; ir++) {
    if (ir == 0) {
      Ar_inner = Ar_outer = PI_D * drdz;
      Az = dr2 / 4 * PI_D;
    } else {
      Ar_inner = PI_D * (2 * ir - 1) * drdz;
      Ar_outer = PI_D * (2 * ir + 1) * drdz;
      Az = PI_D * dr2 / 4 * (isqr(2 * ir + 1) - isqr(2 * ir - 1));
    }
    volume[ir] = Az * dz;
    for (iz = 0; iz < nz; iz++) {
      A(f, f->Q, iz, ir) = 0;
      if (ir == 0)
        dQ = (A(f, f->Er, iz, ir + 1) * Ar_outer + (A(f, f->Ez, iz + 1, ir) - A(f, f->Ez, iz, ir)) * Az) * eps;
      else
        dQ = (A(f, f->Er, iz, ir + 1) * Ar_outer - A(f, f->Er, iz, ir) * Ar_inner +
              (A(f, f->Ez, iz + 1, ir) - A(f, f->Ez, iz, ir)) * Az) *
             eps;
      if (A(f, f->metal_flags, iz, ir) & OFL_IS_METAL || (iz == 0 && f->left_bc == O_NEUMANN) ||
          (iz == nz - 1 && f->right_bc == O_NEUMANN) || (ir == nr - 1 && f->upper_bc == O_NEUMANN)) {
        if (iz == 0)
          dQ += A(f, f->Ez, iz, ir) * Az * eps;
        if (iz == nz - 1)
          dQ -= A(f, f->Ez, iz + 1, ir) * Az * eps;
        if (ir == nr - 1) {
          if (first_pass)
            dQ += A(f, f->Ez, iz, ir + 1) * Ar_outer * eps;
          else
            dQ -= A(f, f->Ez, iz, ir + 1) * Ar_outer * eps;
        }
        *Qsurface += dQ;
      } else {
        A(f, f->Q, iz, ir) = dQ;
        if (first_pass)
          *centroid += dQ * (iz * dz + f->zmin);
      }
    }
  }
}

static double remove_particle_charge(ofields *f, const obeam *b) {
  int64_t ip, iz, ir, nz = f->nz, nr = f->nr;
  int longit_smear = (f->modes & OFL_LONGIT_SMEAR) != 0, radial_smear = (f->modes & OFL_RADIAL_SMEAR) != 0;
  double fa, fb, fl, fr, Qreal = 0, dz = f->dz, dr = f->dr;
  const int16_t *mf = f->metal_flags;
  int lN = f->left_bc == O_NEUMANN, rN = f->right_bc == O_NEUMANN, uN = f->upper_bc == O_NEUMANN;
  for (ip = 0; ip < b->np; ip++) {
    if (!(b->status[ip] & OPART_ACTIVE))
      continue;
    if (longit_smear)
      iz = (b->z[ip] - f->zmin) / dz;
    else
      iz = (b->z[ip] - f->zmin) / dz + 0.5;
    if (radial_smear)
      ir = b->r[ip] / dr;
    else
      ir = b->r[ip] / dr + 0.5;
    if (iz < 0 || iz >= nz)
      continue;
    if (ir < 0 || ir >= nr)
      continue;
    if (radial_smear && ir != nr - 1)
      fb = 1 - (fa = b->r[ip] / dr - ir);
    else {
      fb = 1;
      fa = 0;
    }
    if (longit_smear && iz != nz - 1)
      fl = 1 - (fr = (b->z[ip] - f->zmin) / dz - iz);
    else {
      fl = 1;
      fr = 0;
    }
    if (!(mf[IDX(f, iz, ir)] & OFL_IS_METAL || (iz == 0 && lN) || (iz >= nz - 1 && rN) || (ir >= nr - 1 && uN))) {
      A(f, f->Q, iz, ir) -= b->Q[ip] * fb * fl;
      Qreal += b->Q[ip] * fb * fl;
    }
    if (!(mf[IDX(f, iz + 1, ir)] & OFL_IS_METAL || (iz == 0 && lN) || (iz >= nz && rN) || (ir >= nr - 1 && uN))) {
      A(f, f->Q, iz + 1, ir) -= b->Q[ip] * fb * fr;
      Qreal += b->Q[ip] * fb * fr;
    }
    if (!(mf[IDX(f, iz, ir + 1)] & OFL_IS_METAL || (iz == 0 && lN) || (iz >= nz - 1 && rN) || (ir >= nr && uN))) {
      A(f, f->Q, iz, ir + 1) -= b->Q[ip] * fa * fl;
      Qreal += b->Q[ip] * fa * fl;
    }
    if (!(mf[IDX(f, iz + 1, ir + 1)] & OFL_IS_METAL || (iz == 0 && lN) || (iz >= nz && rN) || (ir >= nr && uN))) {
      A(f, f->Q, iz + 1, ir + 1) -= b->Q[ip] * fa * fr;
      Qreal += b->Q[ip] * fa * fr;
    }
  }
  return Qreal;
}

/* src/poisson_correction.c:564-615 */
static void poisson_guess(ofields *f, double charge, double centroid, int64_t code) {
  int64_t iz, ir, nz = f->nz, nr = f->nr;
  double dist, guess, dz = f->dz, dr = f->dr, zmin = f->zmin;
  switch (code) {
  case 1:
    for (ir = 0; ir < nr; ir++) {
      guess = charge / (2 * PI_D * f->epsilon_o) * log((ir ? ir : 0.5) * dr);
      for (iz = 0; iz < nz; iz++)
        A(f, f->Psi, iz, ir) = guess;
    }
    break;
  case 2:
    for (ir = 0; ir < nr; ir++)
      for (iz = 0; iz < nz; iz++) {
        dist = sqrt(sqr((iz * dz + zmin) - centroid) + sqr(dr * (ir + 0.5)));
        A(f, f->Psi, iz, ir) = charge / (4 * PI_D * f->epsilon_o * dist);
      }
    break;
  case 3:
    for (iz = 0; iz < nz; iz++)
      for (ir = 0; ir < nr; ir++)
        A(f, f->Psi, iz, ir) = 0;
    break;
  default:
    break;
  }
}

/* src/poisson_correction.c:73-454.  Returns 0 ok, nonzero on a bomb() condition. */
int oracle_add_poisson_correction(ofields *f, const obeam *b, int32_t space_charge, opoisson *pc) {
  int64_t iz, ir, nz = f->nz, nr = f->nr;
  double dz = f->dz, dr = f->dr, Qerror, Qreal, Qsurface = 0, centroid = 0;
  double *volume;
  pc->last_iterations = -2;
  if (pc->maximum_iterations <= 0)
    return 0;
  if (pc->pc_start_step == -1) {
    if (f->dti <= 0)
      return 1;
    pc->pc_start_step = (pc->start_time - f->start_time) / f->dti;
  }
  if (pc->step_interval <= 0 || f->time_step % pc->step_interval != 0 || pc->pc_start_step > f->time_step)
    return 0;
  volume = malloc(sizeof(double) * (nr + 1));
  gauss_law_charge(f, volume, &Qsurface, &centroid, 1);
  Qreal = space_charge ? remove_particle_charge(f, b) : 0;
  Qerror = 0;
  for (iz = 0; iz < nz; iz++)
    for (ir = 0; ir < nr; ir++)
      Qerror += A(f, f->Q, iz, ir);
  pc->Qreal = Qreal;
  pc->Qerror = Qerror;
  pc->Qsurface = Qsurface;
  if ((pc->error_charge_threshold > 0 && fabs(Qerror) < pc->error_charge_threshold) ||
      (pc->error_charge_threshold < 0 && (Qreal == 0 || fabs(Qerror / Qreal) < fabs(pc->error_charge_threshold)))) {
    free(volume);
    return 0;
  }
  for (ir = 0; ir < nr; ir++)
    for (iz = 0; iz < nz; iz++)
      A(f, f->Q, iz, ir) = -A(f, f->Q, iz, ir) / (volume[ir] * f->epsilon_o);
  poisson_guess(f, Qerror, centroid, pc->guess_code);
  pc->last_iterations = oracle_solve_poisson_cyl(f->Psi, f->Q, f->metal_flags, f->left_bc, f->right_bc, nz, nr,
                                                 dz, dr, pc->accuracy, pc->maximum_iterations, NULL);
  /* -- Ez, :289-308 */
  for (ir = 0; ir < nr; ir++) {
    for (iz = 1; iz < nz; iz++)
      if (!(A(f, f->metal_flags, iz, ir) & OFL_EZ_ZERO))
        A(f, f->Ez, iz, ir) -= -(A(f, f->Psi, iz, ir) - A(f, f->Psi, iz - 1, ir)) / dz;
    if (f->left_bc == O_NEUMANN)
      A(f, f->Ez, 0, ir) = A(f, f->Ez, 1, ir);
    else
      A(f, f->Ez, 0, ir) = -A(f, f->Ez, 1, ir);
    if (f->right_bc == O_NEUMANN)
      A(f, f->Ez, nz, ir) = A(f, f->Ez, nz - 1, ir);
    else
      A(f, f->Ez, nz, ir) = -A(f, f->Ez, nz - 1, ir);
  }
  for (iz = 0; iz <= nz; iz++)
    A(f, f->Ez, iz, nr - 1) = 0;
  /* -- Er, :310-326 */
  for (ir = 1; ir <= nr; ir++)
    for (iz = 0; iz < nz; iz++)
      if (!(A(f, f->metal_flags, iz, ir) & OFL_ER_ZERO))
        A(f, f->Er, iz, ir) -= -(A(f, f->Psi, iz, ir) - A(f, f->Psi, iz, ir - 1)) / dr;
  for (iz = 0; iz < nz; iz++)
    A(f, f->Er, iz, 0) = -A(f, f->Er, iz, 1);
  if (f->left_bc == O_NEUMANN)
    for (ir = 0; ir <= nr; ir++)
      A(f, f->Er, 0, ir) = 0;
  if (f->right_bc == O_NEUMANN)
    for (ir = 0; ir <= nr; ir++)
      A(f, f->Er, nz - 1, ir) = 0;
  /* recompute Q for output, :332-439 */
  Qsurface = 0;
  gauss_law_charge(f, volume, &Qsurface, &centroid, 0);
  if (space_charge)
    remove_particle_charge(f, b);
  free(volume);
  return 0;
}

/* src/poisson_correction.c:456-562; *last_field_factor plays the role of the static at :463. */
int oracle_compute_imposed_efield(ofields *f, int32_t initialize, opoisson *pc, double *last_field_factor) {
  int64_t iz, ir, nz = f->nz, nr = f->nr;
  double dr = f->dr, dz = f->dz, field_factor;
  if (pc->maximum_iterations <= 0)
    return 0;
  if (pc->pc_start_step == -1) {
    if (f->dti <= 0)
      return 1;
    pc->pc_start_step = (pc->start_time - f->start_time) / f->dti;
  }
  if (initialize)
    pc->last_iterations = oracle_solve_poisson_cyl(f->PsiImposed, NULL, f->metal_flags, f->left_bc, f->right_bc,
                                                   nz, nr, dz, dr, pc->accuracy, pc->maximum_iterations, NULL);
  field_factor = 1;
  if (f->imposedEFieldRampTime) {
    double T = f->imposedEFieldRampTime, F = f->imposedEFieldFlatTopTime;
    if (f->time < T)
      field_factor = f->time / T;
    else if (f->time < (T + F))
      field_factor = 1;
    else if (f->time < (2 * T + F))
      field_factor = ((2 * T + F) - f->time) / T;
    else
      field_factor = 0;
    if (field_factor < 0)
      field_factor = 0;
  }
  if (field_factor == *last_field_factor)
    return 0;
  *last_field_factor = field_factor;
  for (ir = 0; ir < nr; ir++) {
    for (iz = 1; iz < nz; iz++)
      if (!(A(f, f->metal_flags, iz, ir) & OFL_EZ_ZERO))
        A(f, f->EzImposed, iz, ir) = -field_factor * (A(f, f->PsiImposed, iz, ir) - A(f, f->PsiImposed, iz - 1, ir)) / dz;
    if (f->left_bc == O_NEUMANN)
      A(f, f->EzImposed, 0, ir) = A(f, f->EzImposed, 1, ir);
    else
      A(f, f->EzImposed, 0, ir) = -A(f, f->EzImposed, 1, ir);
    if (f->right_bc == O_NEUMANN)
      A(f, f->EzImposed, nz, ir) = A(f, f->EzImposed, nz - 1, ir);
    else
      A(f, f->EzImposed, nz, ir) = -A(f, f->EzImposed, nz - 1, ir);
  }
  for (iz = 0; iz <= nz; iz++)
    A(f, f->EzImposed, iz, nr - 1) = 0;
  for (ir = 1; ir <= nr; ir++)
    for (iz = 0; iz < nz; iz++)
      if (!(A(f, f->metal_flags, iz, ir) & OFL_ER_ZERO))
        A(f, f->ErImposed, iz, ir) = -field_factor * (A(f, f->PsiImposed, iz, ir) - A(f, f->PsiImposed, iz, ir - 1)) / dr;
  for (iz = 0; iz < nz; iz++)
    A(f, f->ErImposed, iz, 0) = -A(f, f->ErImposed, iz, 1);
  if (f->left_bc == O_NEUMANN)
    for (ir = 0; ir <= nr; ir++)
      A(f, f->ErImposed, 0, ir) = 0;
  if (f->right_bc == O_NEUMANN)
    for (ir = 0; ir <= nr; ir++)
      A(f, f->ErImposed, nz - 1, ir) = 0;
  return 0;
}

/* One step of the loop body of src/integrate.c:134-194 restricted to the hot path. */
int oracle_step(ofields *f, obeam *b, const oscreen *screens, int32_t n_screens, int32_t compat,
                opoisson *pc, int32_t n_lines, const int32_t *line_dir, const int64_t *line_pos,
                const int64_t *line_start, const int64_t *line_end, const double *line_J,
                oscreen_row *rows, int64_t max_rows, int64_t *n_rows,
                olost_row *lost, int64_t max_lost, int64_t *n_lost) {
  int32_t k;
  int rc = 0;
  f->time = f->time_step * f->dti + f->start_time;
  oracle_zero_current_arrays(f);
  if (f->modes & OFL_SPACE_CHARGE)
    oracle_add_particle_currents(f, b);
  for (k = 0; k < n_lines; k++)
    oracle_add_line_current(f, line_dir[k], line_pos[k], line_start[k], line_end[k], line_J[k]);
  if (pc)
    rc = oracle_add_poisson_correction(f, b, (f->modes & OFL_SPACE_CHARGE) != 0, pc);
  if (rc)
    return rc;
  oracle_advance_field_solutions(f);
  rc = oracle_advance_particles(f, b, screens, n_screens, -1, compat, rows, max_rows, n_rows, lost, max_lost, n_lost);
  f->time_step++;
  return rc;
}
