/*
 * FP64 restatement of the reference's Ewald direct-space sum with the
 * self-energy term, CpuEwaldDirect.cpp:31-131 -- TEST INFRASTRUCTURE (the
 * checker of the GPU direct-space kernel; loaded only by tests/).
 *
 * Loop order, the minimum-image while loops, the strict cutoff r2 < rcut2
 * and the operand order of every product are kept, so that this function
 * reproduces the reference's compiled code bit for bit (the tests compare it
 * with oracle/_ref where that builds).  As in the reference, the outputs
 * fx, fy, fz are the energy gradient dE/dr_i (a pair term adds f*d to atom
 * i and subtracts it from atom j, with f = -qq(2/sqrt(pi) a e^{-r^2/2s^2} +
 * erfc(a r)/r)/r^2, a = 1/(sqrt(2) sigma), d = r_i - r_j), and excluded
 * pairs are skipped without a correction term.
 *
 * Exclusions follow setupExclusions (:31-40): iblo14[j] is the running end
 * of atom j's list in inb14 (1-based atom numbers, sign ignored); pair
 * (j, i > j) is skipped when i is in j's list.  n_iblo14 = 0: no exclusions.
 */
#include <math.h>
#include <stdlib.h>

#include "gse_oracle.h"

static const double PI_E = 3.14159265358979323846; /* TypesGSE.h pi_dbl */

double orc_ewald_direct(double sigma, double rcut, double boxx, double boxy, double boxz,
                        double ccelec, int n, const double *xyzq, int n_iblo14,
                        const int *iblo14_in, const int *inb14_in, double *fx, double *fy,
                        double *fz) {
  int *inb14 = NULL;
  if (n_iblo14 > 0) {
    const int n_inb14 = iblo14_in[n_iblo14 - 1];
    inb14 = (int *)malloc(sizeof(int) * (size_t)(n_inb14 > 0 ? n_inb14 : 1));
    for (int i = 0; i < n_inb14; i++) inb14[i] = abs(inb14_in[i]) - 1;
  }
  const double rcut2 = rcut * rcut;
  const double hboxx = 0.5 * boxx, hboxy = 0.5 * boxy, hboxz = 0.5 * boxz;
  const double inv_sqrt2sigma = 1.0 / (sqrt(2.0) * sigma);
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double two_sqrtpi = 2.0 / sqrt(PI_E);
  for (int i = 0; i < n; i++) {
    fx[i] = 0.0;
    fy[i] = 0.0;
    fz[i] = 0.0;
  }
  double E = 0.0;
  for (int j = 0; j < n - 1; j++) {
    const double xj = xyzq[4 * j], yj = xyzq[4 * j + 1], zj = xyzq[4 * j + 2];
    const double qj = xyzq[4 * j + 3] * ccelec;
    int excl_start = 0, excl_end = 0;
    if (n_iblo14 > 0) {
      excl_start = (j >= 1) ? iblo14_in[j - 1] : 0;
      excl_end = iblo14_in[j];
    }
    for (int i = j + 1; i < n; i++) {
      int exclude = 0;
      for (int k = excl_start; k < excl_end; k++)
        if (i == inb14[k]) {
          exclude = 1;
          break;
        }
      if (exclude) continue;
      double dx = xyzq[4 * i] - xj, dy = xyzq[4 * i + 1] - yj, dz = xyzq[4 * i + 2] - zj;
      const double qi = xyzq[4 * i + 3];
      while (dx < -hboxx) dx += boxx;
      while (dx > hboxx) dx -= boxx;
      while (dy < -hboxy) dy += boxy;
      while (dy > hboxy) dy -= boxy;
      while (dz < -hboxz) dz += boxz;
      while (dz > hboxz) dz -= boxz;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < rcut2) {
        const double qq = qj * qi;
        const double r = sqrt(r2);
        const double erfcval = erfc(r * inv_sqrt2sigma);
        const double expval = exp(-r2 * inv_2sigmasq);
        const double f = -qq * (two_sqrtpi * inv_sqrt2sigma * expval + erfcval / r) / r2;
        const double fijx = f * dx, fijy = f * dy, fijz = f * dz;
        fx[i] += fijx;
        fy[i] += fijy;
        fz[i] += fijz;
        fx[j] -= fijx;
        fy[j] -= fijy;
        fz[j] -= fijz;
        E += qq * erfcval / r;
      }
    }
  }
  double Eself = 0.0;
  for (int i = 0; i < n; i++) Eself += xyzq[4 * i + 3] * xyzq[4 * i + 3];
  Eself *= -ccelec / sqrt(2.0 * PI_E * sigma * sigma);
  free(inb14);
  return E + Eself;
}
