/*
 * Drives every function of the C oracle once on small inputs, for a build
 * under -fsanitize=address,undefined (TEST INFRASTRUCTURE: built and run by
 * tests/test_oracle_sanitize.py on the CPU, never loaded by the product).
 *
 * Besides running each routine under the sanitizers it checks relations
 * that hold whatever the inputs are, so that a routine that silently reads
 * the wrong memory also fails here:
 *   - the k-space (Parseval) and real-space energies of a step agree;
 *   - forward then backward FFT returns the input times X*Y*Z;
 *   - Newton's third law: the direct-space forces sum to ~0;
 *   - the SPME force is minus the central difference of the SPME energy;
 *   - the exact Mehrstellen solve is a fixed point of the relaxation.
 * Exit status 0 when every check passes; the first failure is printed.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gse_oracle.h"

static unsigned long long rng_state = 88172645463325252ull;

static double urand(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int fails = 0;

static void check(int ok, const char *what, double a, double b) {
  if (!ok) {
    fprintf(stderr, "FAIL %s: %.17g vs %.17g\n", what, a, b);
    fails++;
  }
}

static double *dalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

int main(void) {
  const int X = 12, Y = 10, Z = 9, n = 24;
  const double bx = 12.5, by = 10.0, bz = 9.5, ccelec = 332.0716;
  const size_t ng = (size_t)X * Y * Z, nk = (size_t)(X / 2 + 1) * Y * Z;
  double *xyzq = dalloc(4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    xyzq[4 * i] = bx * urand();
    xyzq[4 * i + 1] = by * urand();
    xyzq[4 * i + 2] = bz * urand();
    xyzq[4 * i + 3] = (i % 2) ? -0.8 : 0.8;
  }

  /* FFT round trip */
  double *g = dalloc(ng), *g2 = dalloc(ng), *k = dalloc(2 * nk);
  for (size_t i = 0; i < ng; i++) g[i] = urand() - 0.5;
  orc_fft_r2c(X, Y, Z, g, k);
  orc_fft_c2r(X, Y, Z, k, g2);
  double err = 0.0;
  for (size_t i = 0; i < ng; i++) err = fmax(err, fabs(g2[i] / (double)ng - g[i]));
  check(err < 1e-13, "fft round trip", err, 0.0);
  double c1[16];
  for (int i = 0; i < 16; i++) c1[i] = urand();
  orc_fft_1d(8, c1, -1);
  orc_fft_1d(8, c1, 1);

  /* GSE-k step, both modes, every output */
  orc_gse_params p = {X, Y, Z, 2.1, 0.4, 3.0, 3.0, bx, by, bz, ccelec};
  double s1, s2, r1, r2;
  orc_gse_derive(&p, &s1, &s2, &r1, &r2);
  double *f = dalloc(9 * (size_t)n), *pot = dalloc(n), vir[9], e, ek;
  double *rho = dalloc(ng), *phi = dalloc(ng);
  for (int mode = 0; mode < 2; mode++) {
    orc_gse_step(&p, mode, n, xyzq, &e, &ek, f, f + n, f + 2 * n, f + 3 * n, f + 4 * n,
                 f + 5 * n, pot, vir, rho, phi);
    check(fabs(e - ek) <= 1e-9 * fabs(e), "Parseval vs real-space energy", e, ek);
  }
  check(orc_support_size(X, Y, Z, r1, bx, by, bz) > 0, "support size", 0, 0);
  orc_spread(X, Y, Z, s1, r1, n, xyzq, bx, by, bz, rho, 0);
  orc_spread_on_grid(X, Y, Z, s2, r2, bx, by, bz, rho, phi);
  double kx[7] = {0.1, 0.2, 0.5, 1.0, 0.5, 0.2, 0.1};
  orc_conv_run_x(X, Y, Z, 3, kx, rho, phi);
  orc_conv_run_xyz(X, Y, Z, 3, 2, 1, kx, kx + 1, kx + 2, rho, phi);
  orc_interpolate(X, Y, Z, s1, r1, n, xyzq, bx, by, bz, phi, f, f + n, f + 2 * n);
  orc_gather_cube(X, Y, Z, s1, r1, n, xyzq, bx, by, bz, phi, pot, f, f + n, f + 2 * n, vir);
  orc_fft_r2c(X, Y, Z, rho, k);
  for (int kind = 0; kind < 2; kind++) {
    orc_greens_mul(kind, X, Y, Z, bx, by, bz, s2, k);
    orc_fft_solve(kind, X, Y, Z, s2, bx, by, bz, rho, phi);
  }
  orc_fft_r2c(X, Y, Z, rho, k);
  (void)orc_energy_kspace(0, X, Y, Z, ccelec, bx, by, bz, s2, k, vir);
  (void)orc_energy(X, Y, Z, ccelec, bx, by, bz, rho, phi);
  orc_efield(X, Y, Z, bx, by, bz, phi, g, g2, rho);
  double dip[3];
  orc_dipole_sum(X, Y, Z, bx, by, bz, rho, dip);
  double es, en;
  orc_corrections(2.1, bx, by, bz, ccelec, n, xyzq, &es, &en);
  (void)orc_ewald_recip(2.1, 4, bx, by, bz, ccelec, n, xyzq, f, f + n, f + 2 * n);

  /* Ewald direct with and without exclusions; forces sum to zero */
  int *iblo = (int *)calloc((size_t)n, sizeof(int)), *inb = (int *)calloc((size_t)n, sizeof(int));
  for (int j = 0; j < n; j++) {
    iblo[j] = (j < n - 1) ? j + 1 : n - 1; /* atom j excludes atom j+1 */
    if (j < n - 1) inb[j] = j + 2;
  }
  for (int ex = 0; ex < 2; ex++) {
    (void)orc_ewald_direct(2.1, 4.5, bx, by, bz, ccelec, n, xyzq, ex ? n : 0, iblo, inb, f,
                           f + n, f + 2 * n);
    double sx = 0, sy = 0, sz = 0, mx = 0;
    for (int i = 0; i < n; i++) {
      sx += f[i];
      sy += f[n + i];
      sz += f[2 * n + i];
      mx = fmax(mx, fabs(f[i]));
    }
    check(fabs(sx) + fabs(sy) + fabs(sz) <= 1e-12 * (mx + 1.0), "direct force sum", sx, 0);
  }

  /* SPME: orders 4, 6, 8 */
  double th[8], dth[8], *pre = dalloc(16);
  for (int order = 4; order <= 8; order += 2) {
    orc_pme_theta(order, 0.37, th, dth);
    double s = 0.0;
    for (int i = 0; i < order; i++) s += th[i];
    check(fabs(s - 1.0) < 1e-14, "B-spline partition of unity", s, 1.0);
    orc_pme_prefac(order, 16, pre);
    orc_pme_spread(X, Y, Z, order, n, xyzq, bx, by, bz, rho);
    orc_fft_r2c(X, Y, Z, rho, k);
    orc_pme_scalar_sum(X, Y, Z, order, bx, by, bz, 0.3, ccelec, k, &e, vir);
    orc_pme_gather(X, Y, Z, order, n, xyzq, bx, by, bz, ccelec, phi, f, f + n, f + 2 * n);
    orc_pme_step(X, Y, Z, order, n, xyzq, bx, by, bz, 0.3, ccelec, &e, vir, f, f + n, f + 2 * n,
                 rho);
    /* the force is minus the exact gradient of the SPME energy: central
     * difference along x of atom 3 */
    const double fx3 = f[3], hstep = 1e-4, x3 = xyzq[12];
    double ep, em;
    xyzq[12] = x3 + hstep;
    orc_pme_step(X, Y, Z, order, n, xyzq, bx, by, bz, 0.3, ccelec, &ep, NULL, NULL, NULL, NULL,
                 NULL);
    xyzq[12] = x3 - hstep;
    orc_pme_step(X, Y, Z, order, n, xyzq, bx, by, bz, 0.3, ccelec, &em, NULL, NULL, NULL, NULL,
                 NULL);
    xyzq[12] = x3;
    const double fd = -(ep - em) / (2.0 * hstep);
    check(fabs(fd - fx3) <= 1e-6 * fabs(fx3) + 1e-9, "SPME force vs energy difference", fx3, fd);
  }

  /* GSE-r: the exact solve is a fixed point of the relaxation; GSE-u runs */
  double *phi2 = dalloc(ng);
  orc_spread(X, Y, Z, s1, r1, n, xyzq, bx, by, bz, rho, 0);
  double mean = 0.0;
  for (size_t i = 0; i < ng; i++) mean += rho[i];
  mean /= (double)ng;
  for (size_t i = 0; i < ng; i++) rho[i] -= mean;
  orc_mehrstellen_solve(X, Y, Z, bx, by, bz, rho, phi);
  for (size_t i = 0; i < ng; i++) g[i] = -4.0 * 3.14159265358979323846 * rho[i];
  memcpy(phi2, phi, sizeof(double) * ng);
  orc_rbgs_hermitian(X, Y, Z, bx, by, bz, g, phi2, 2);
  err = 0.0;
  double ref = 0.0;
  for (size_t i = 0; i < ng; i++) {
    err = fmax(err, fabs(phi2[i] - phi[i]));
    ref = fmax(ref, fabs(phi[i]));
  }
  check(err <= 1e-10 * ref, "Mehrstellen fixed point", err, ref);
  (void)orc_mehrstellen_g(X, Y, Z, bx, by, bz, 1, 2, 3);
  orc_gser_step(&p, 3, n, xyzq, &e, f, f + n, f + 2 * n, rho, g, phi);
  orc_gser_step(&p, 0, n, xyzq, &e, f, f + n, f + 2 * n, NULL, NULL, NULL);
  orc_gseu_step(3, 2.0, &p, n, xyzq, &e, f, f + n, f + 2 * n, phi);

  free(xyzq);
  free(g);
  free(g2);
  free(k);
  free(f);
  free(pot);
  free(rho);
  free(phi);
  free(phi2);
  free(pre);
  free(iblo);
  free(inb);
  if (fails) return 1;
  printf("sanitized oracle run: all checks passed\n");
  return 0;
}
