/*
 * FP64 restatements of the reference's GSE-u and GSE-r pipelines --
 * TEST INFRASTRUCTURE (checkers of the GPU CpuGSEu / CpuGSEr; loaded only
 * by tests/).
 *
 * GSE-u (CpuGSEu.cpp:40-92): sigma1 spread into rhoS; for j = 0..N-1,
 *   sigma_j = sqrt(sigma^2 b^(2j) - 2 sigma1^2), rhoG_j = the on-grid
 *   sigma_j convolution of rhoS (radius from lambdaSigma1 sqrt2 sigma_j)
 *   scaled by (2 pi sigma_j^2)^(3/2) / sqrt(sigma_j^2); phiM = sum of the
 *   rhoG_j added from j = N-1 down to 0; E = 2 ln(b) * 1/2 ccelec h^3 sum
 *   rhoS phiM; forces = 2 ln(b) ccelec * interpolateForce(sigma1, phiM).
 *   The reference also writes rhoG%d.txt at every solve; that is not kept.
 *
 * GSE-r (CpuGSEr.cpp:127-160 with CpuMultiGridSolver.cpp:59-118, 204-290):
 *   sigma1 spread into rho; sigma2 on-grid convolution into rhoS; phiM
 *   solves the HERMITIAN (Mehrstellen) discretisation A phi = B rho' with
 *   rho' = -4 pi rhoS, A the 19-point operator and B the 7-point source
 *   average of relaxRBGS; E = 1/2 ccelec h^3 sum rho phiM; forces =
 *   ccelec * interpolateForce(sigma1, phiM).  The reference relaxes 200
 *   red-black Gauss-Seidel sweeps from phi = 0 (one level, numLevel = 1);
 *   orc_rbgs_hermitian restates those sweeps point for point.
 *   orc_mehrstellen_solve gives the exact solution of the same linear system
 *   in k-space (zero mean): phi^(k) = -4 pi B^(k)/A^(k) rhoS^(k), the
 *   symbols written with s_d = 1 - cos(2 pi m_d / M_d) = 2 sin^2(pi m_d / M_d):
 *     A^ = -2 sum_d b_d s_d - 4 sum_{d<e} c_de (s_d + s_e - s_d s_e)
 *     B^ = -1 + (s_x + s_y + s_z)/6
 *   with b_d = -5/6 h_d^-2 + S/6, c_de = -(h_d^-2 + h_e^-2)/12, S = sum h^-2
 *   (the coefficients of CpuMultiGridSolver.cpp:237-251).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gse_oracle.h"

static const double PI_R = 3.14159265358979323846;

static int wr(int i, int m) {
  int r = i % m;
  return r < 0 ? r + m : r;
}

/* phi at a wrapped position, as Grid::getPos (Grid.h:32-37) */
#define AT(g, x, y, z) (g)[wr(x, X) + ((size_t)wr(y, Y) + (size_t)wr(z, Z) * Y) * X]

void orc_rbgs_hermitian(int X, int Y, int Z, double bx, double by, double bz,
                        const double *rho, double *phi, int nsweep) {
  double ihx2 = (double)X / bx, ihy2 = (double)Y / by, ihz2 = (double)Z / bz;
  ihx2 *= ihx2;
  ihy2 *= ihy2;
  ihz2 *= ihz2;
  const double sum_inv = ihx2 + ihy2 + ihz2;
  const double a_A = 4.0 / 3.0 * sum_inv;
  const double inv_a_A = 1.0 / a_A;
  const double a_B = -0.5, b_B = -1.0 / 12.0;
  const double bx_A = -5.0 / 6.0 * ihx2 + sum_inv / 6.0;
  const double by_A = -5.0 / 6.0 * ihy2 + sum_inv / 6.0;
  const double bz_A = -5.0 / 6.0 * ihz2 + sum_inv / 6.0;
  const double cxy_A = -1.0 / 12.0 * (ihx2 + ihy2);
  const double cxz_A = -1.0 / 12.0 * (ihx2 + ihz2);
  const double cyz_A = -1.0 / 12.0 * (ihy2 + ihz2);
  for (int sweep = 0; sweep < nsweep; sweep++) {
    int iz0 = 0;
    for (int ipass = 0; ipass < 2; ipass++) {
      int iy0 = iz0;
      for (int iz = 0; iz < Z; iz++) {
        int ix0 = iy0;
        for (int iy = 0; iy < Y; iy++) {
          for (int ix = ix0; ix < X; ix += 2) {
            const size_t p = ix + ((size_t)iy + (size_t)iz * Y) * X;
            const double B_rho = a_B * AT(rho, ix, iy, iz) +
                                 b_B * (AT(rho, ix - 1, iy, iz) + AT(rho, ix + 1, iy, iz)) +
                                 b_B * (AT(rho, ix, iy - 1, iz) + AT(rho, ix, iy + 1, iz)) +
                                 b_B * (AT(rho, ix, iy, iz - 1) + AT(rho, ix, iy, iz + 1));
            const double bA_phi = bx_A * (AT(phi, ix - 1, iy, iz) + AT(phi, ix + 1, iy, iz)) +
                                  by_A * (AT(phi, ix, iy - 1, iz) + AT(phi, ix, iy + 1, iz)) +
                                  bz_A * (AT(phi, ix, iy, iz - 1) + AT(phi, ix, iy, iz + 1));
            const double cA_phi =
                cxy_A * (AT(phi, ix - 1, iy - 1, iz) + AT(phi, ix - 1, iy + 1, iz) +
                         AT(phi, ix + 1, iy - 1, iz) + AT(phi, ix + 1, iy + 1, iz)) +
                cxz_A * (AT(phi, ix - 1, iy, iz - 1) + AT(phi, ix - 1, iy, iz + 1) +
                         AT(phi, ix + 1, iy, iz - 1) + AT(phi, ix + 1, iy, iz + 1)) +
                cyz_A * (AT(phi, ix, iy - 1, iz - 1) + AT(phi, ix, iy - 1, iz + 1) +
                         AT(phi, ix, iy + 1, iz - 1) + AT(phi, ix, iy + 1, iz + 1));
            phi[p] = inv_a_A * (B_rho - bA_phi - cA_phi);
          }
          ix0 = (ix0 + 1) % 2;
        }
        iy0 = (iy0 + 1) % 2;
      }
      iz0 = (iz0 + 1) % 2;
    }
  }
}

/* the k-space factor -4 pi B^/(A^ XYZ) of the Mehrstellen solve (0 at k=0) */
double orc_mehrstellen_g(int X, int Y, int Z, double bx, double by, double bz, int ix, int iy,
                         int iz) {
  if (ix == 0 && iy == 0 && iz == 0) return 0.0;
  double ih[3] = {(double)X / bx, (double)Y / by, (double)Z / bz};
  for (int d = 0; d < 3; d++) ih[d] *= ih[d];
  const double S = ih[0] + ih[1] + ih[2];
  const double sx = 2.0 * pow(sin(PI_R * ix / X), 2.0), sy = 2.0 * pow(sin(PI_R * iy / Y), 2.0),
               sz = 2.0 * pow(sin(PI_R * iz / Z), 2.0);
  const double bxA = -5.0 / 6.0 * ih[0] + S / 6.0, byA = -5.0 / 6.0 * ih[1] + S / 6.0,
               bzA = -5.0 / 6.0 * ih[2] + S / 6.0;
  const double cxy = -(ih[0] + ih[1]) / 12.0, cxz = -(ih[0] + ih[2]) / 12.0,
               cyz = -(ih[1] + ih[2]) / 12.0;
  const double A = -2.0 * (bxA * sx + byA * sy + bzA * sz) -
                   4.0 * (cxy * (sx + sy - sx * sy) + cxz * (sx + sz - sx * sz) +
                          cyz * (sy + sz - sy * sz));
  const double B = -1.0 + (sx + sy + sz) / 6.0;
  return -4.0 * PI_R * B / (A * (double)X * Y * Z);
}

/* phi = the zero-mean solution of A phi = B (-4 pi rhoS) */
void orc_mehrstellen_solve(int X, int Y, int Z, double bx, double by, double bz,
                           const double *rhoS, double *phi) {
  const int Xh = X / 2 + 1;
  double *k = (double *)malloc(sizeof(double) * 2 * (size_t)Xh * Y * Z);
  orc_fft_r2c(X, Y, Z, rhoS, k);
  for (int iz = 0; iz < Z; iz++)
    for (int iy = 0; iy < Y; iy++)
      for (int ix = 0; ix < Xh; ix++) {
        const double g = orc_mehrstellen_g(X, Y, Z, bx, by, bz, ix, iy, iz);
        double *c = k + 2 * (ix + ((size_t)iy + (size_t)iz * Y) * Xh);
        c[0] *= g;
        c[1] *= g;
      }
  orc_fft_c2r(X, Y, Z, k, phi);
  free(k);
}

void orc_gser_step(const orc_gse_params *p, int nsweep, int n, const double *xyzq,
                   double *energy, double *fx, double *fy, double *fz, double *rho_out,
                   double *rhoS_out, double *phi_out) {
  const int X = p->X, Y = p->Y, Z = p->Z;
  const size_t sz = (size_t)X * Y * Z;
  double s1, s2, rc1, rc2;
  orc_gse_derive(p, &s1, &s2, &rc1, &rc2);
  double *rho = (double *)malloc(sizeof(double) * sz);
  double *rhoS = (double *)malloc(sizeof(double) * sz);
  double *phi = (double *)calloc(sz, sizeof(double));
  orc_spread(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, rho, 0);           /* :128-131 */
  orc_spread_on_grid(X, Y, Z, s2, rc2, p->bx, p->by, p->bz, rho, rhoS);         /* :137-140 */
  if (nsweep > 0) {                                                             /* :146-152 */
    double *rt = (double *)malloc(sizeof(double) * sz);
    for (size_t i = 0; i < sz; i++) rt[i] = rhoS[i] * (-4.0 * PI_R);
    orc_rbgs_hermitian(X, Y, Z, p->bx, p->by, p->bz, rt, phi, nsweep);
    free(rt);
  } else {
    orc_mehrstellen_solve(X, Y, Z, p->bx, p->by, p->bz, rhoS, phi);
  }
  if (energy) *energy = orc_energy(X, Y, Z, p->ccelec, p->bx, p->by, p->bz, rho, phi);
  if (fx || fy || fz) {
    const size_t m = (size_t)(n > 0 ? n : 1);
    double *a = (double *)malloc(sizeof(double) * 3 * m);
    orc_interpolate(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, phi, a, a + m, a + 2 * m);
    for (int i = 0; i < n; i++) {
      if (fx) fx[i] = a[i] * p->ccelec;
      if (fy) fy[i] = a[m + i] * p->ccelec;
      if (fz) fz[i] = a[2 * m + i] * p->ccelec;
    }
    free(a);
  }
  if (rho_out) memcpy(rho_out, rho, sizeof(double) * sz);
  if (rhoS_out) memcpy(rhoS_out, rhoS, sizeof(double) * sz);
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * sz);
  free(rho);
  free(rhoS);
  free(phi);
}

void orc_gseu_step(int N, double b, const orc_gse_params *p, int n, const double *xyzq,
                   double *energy, double *fx, double *fy, double *fz, double *phi_out) {
  const int X = p->X, Y = p->Y, Z = p->Z;
  const size_t sz = (size_t)X * Y * Z;
  double s1, s2, rc1, rc2;
  orc_gse_derive(p, &s1, &s2, &rc1, &rc2);
  double *rhoS = (double *)malloc(sizeof(double) * sz);
  double *phi = (double *)calloc(sz, sizeof(double));
  double *g = (double *)malloc(sizeof(double) * sz * (size_t)N);
  orc_spread(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, rhoS, 0);  /* :40-43 */
  for (int j = 0; j < N; j++) {                                         /* :51-58 */
    const double sj = sqrt(p->sigma * p->sigma * pow(b, 2 * j) - 2 * s1 * s1);
    double *gj = g + sz * (size_t)j;
    orc_spread_on_grid(X, Y, Z, sj, p->lambda_sigma1 * sqrt(2.0) * sj, p->bx, p->by, p->bz,
                       rhoS, gj);
    const double fac = pow(2 * PI_R * sj * sj, 3.0 / 2.0) / pow(sj * sj, 1.0 / 2.0);
    for (size_t i = 0; i < sz; i++) gj[i] *= fac;
  }
  for (int j = N - 1; j >= 0; j--)                                      /* :60-66 */
    for (size_t i = 0; i < sz; i++) phi[i] += g[sz * (size_t)j + i];
  const double lb = 2.0 * log(b);
  if (energy) *energy = lb * orc_energy(X, Y, Z, p->ccelec, p->bx, p->by, p->bz, rhoS, phi);
  if (fx || fy || fz) {                                                 /* :73-84 */
    const size_t m = (size_t)(n > 0 ? n : 1);
    double *a = (double *)malloc(sizeof(double) * 3 * m);
    orc_interpolate(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, phi, a, a + m, a + 2 * m);
    const double pref = lb * p->ccelec;
    for (int i = 0; i < n; i++) {
      if (fx) fx[i] = a[i] * pref;
      if (fy) fy[i] = a[m + i] * pref;
      if (fz) fz[i] = a[2 * m + i] * pref;
    }
    free(a);
  }
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * sz);
  free(rhoS);
  free(phi);
  free(g);
}
