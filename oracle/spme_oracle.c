/*
 * FP64 restatement of the reference's smooth particle-mesh Ewald (SPME)
 * reciprocal path, cuda/CudaPMERecip.cu -- TEST INFRASTRUCTURE (the checker
 * of the GPU SPME comparator; loaded only by tests/).
 *
 * The reference runs this path in FP32 on CUDA only, so it cannot be built
 * here; this file keeps its conventions so that the GPU comparator and this
 * oracle compute the same function:
 *   - fractional grid position (spread :436-588, gather :1612-1700):
 *       w = x/L + 2; fr = M (w - (floor(w + 1/2) - 1/2)); i0 = (int) fr;
 *       the weights theta[j] of grid point i0 + j (mod M), j = 0..order-1,
 *       come from the B-spline recursion at frac = fr - i0
 *       (calc_theta_dtheta :266-327, fill_bspline_host :3411-3443);
 *   - B-spline moduli prefac = 1 / |sum_j bsp_arr[j] e^{2 pi i k j / M}|^2
 *     with the small-modulus fix of dftmod (:3385-3409, calc_prefac
 *     :3448-3478);
 *   - scalar_sum (:946-1256, host :2917-3033): over the r2c half spectrum,
 *     eterm = exp(-pi^2 m^2 / kappa^2) / (pi V) prefac_x prefac_y prefac_z / m^2
 *     with m = (m1/Lx, m2/Ly, m3/Lz), k = 0 dropped; E = ccelec/2 sum of
 *     eterm |Q^|^2 over each element and, for x-index >= 1, its conjugate
 *     partner (so the x Nyquist plane of an even grid enters twice, as in
 *     the reference); virial_ab = -ccelec/2 sum eterm |Q^|^2
 *     (2 (pi^2/kappa^2 + 1/m^2) m_a m_b - delta_ab) over the same terms;
 *     the spectrum is multiplied by eterm in place;
 *   - gather (:1612-1890): dE/dr_a = q ccelec M_a/L_a sum dtheta_a theta theta
 *     Qc over the order^3 points, Qc = c2r of the multiplied spectrum; the
 *     driver negates it into the force (gpu_gse.cu:145-149).
 * Pinned by the tests against the Ewald reciprocal sum (orc_ewald_recip),
 * central finite differences of the energy (forces, virial), the B-spline
 * partition of unity and closed-form cardinal B-spline values.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gse_oracle.h"

static const double PI_S = 3.14159265358979323846;

/* calc_theta_dtheta (:266-327): theta[0..order-1] and dtheta of frac w */
void orc_pme_theta(int order, double w, double *theta, double *dtheta) {
  theta[order - 1] = 0.0;
  theta[1] = w;
  theta[0] = 1.0 - w;
  for (int k = 3; k <= order - 1; k++) {
    const double div = 1.0 / (double)(k - 1);
    theta[k - 1] = div * w * theta[k - 2];
    for (int j = 1; j <= k - 2; j++)
      theta[k - j - 1] = div * ((w + j) * theta[k - j - 2] + (k - j - w) * theta[k - j - 1]);
    theta[0] = div * (1.0 - w) * theta[0];
  }
  dtheta[0] = -theta[0];
  for (int j = 2; j <= order; j++) dtheta[j - 1] = theta[j - 2] - theta[j - 1];
  const double div = 1.0 / (double)(order - 1);
  theta[order - 1] = div * w * theta[order - 2];
  for (int j = 1; j <= order - 2; j++)
    theta[order - j - 1] =
        div * ((w + j) * theta[order - j - 2] + (order - j - w) * theta[order - j - 1]);
  theta[0] = div * (1.0 - w) * theta[0];
}

/* calc_prefac (:3448-3478) with dftmod (:3385-3409) for one axis */
void orc_pme_prefac(int order, int nfft, double *prefac) {
  double th[16], dth[16];
  orc_pme_theta(order, 0.0, th, dth);
  double *arr = (double *)calloc((size_t)nfft + order + 2, sizeof(double));
  for (int i = 2; i <= order + 1; i++) arr[i - 1] = th[i - 2];
  const double r = 2.0 * PI_S / (double)nfft;
  for (int k = 1; k <= nfft; k++) {
    double s1 = 0.0, s2 = 0.0;
    const double a1 = (k - 1) * r;
    for (int j = 1; j < nfft; j++) {
      const double a = a1 * (j - 1);
      s1 += arr[j - 1] * cos(a);
      s2 += arr[j - 1] * sin(a);
    }
    prefac[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= nfft; k++)
    if (prefac[k - 1] < 1.0e-10) {
      /* neighbours of a vanishing modulus (odd orders at the Nyquist
       * frequency only, never an end point) */
      const double lo = prefac[(k - 2 + nfft) % nfft], hi = prefac[k % nfft];
      prefac[k - 1] = 0.5 * (lo + hi);
    }
  for (int k = 0; k < nfft; k++) prefac[k] = 1.0 / prefac[k];
  free(arr);
}

/* fractional grid position of one coordinate (spread :471-476) */
static void pme_pos(double x, double recip, int nfft, int *i0, double *frac) {
  const double w = x * recip + 2.0;
  const double fr = (double)nfft * (w - (floor(w + 0.5) - 0.5));
  int fi = (int)fr;
  *frac = fr - (double)fi;
  *i0 = fi;
}

static int wrapp(int i, int m) {
  int r = i % m;
  return r < 0 ? r + m : r;
}

/* spread_charge_ortho_<order> (:436-588, :701-934): Q = sum q theta theta theta */
void orc_pme_spread(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                    double by, double bz, double *Q) {
  memset(Q, 0, sizeof(double) * (size_t)X * Y * Z);
  double tx[16], ty[16], tz[16], d[16];
  for (int i = 0; i < n; i++) {
    int ix, iy, iz;
    double fx, fy, fz;
    pme_pos(xyzq[4 * i + 0], 1.0 / bx, X, &ix, &fx);
    pme_pos(xyzq[4 * i + 1], 1.0 / by, Y, &iy, &fy);
    pme_pos(xyzq[4 * i + 2], 1.0 / bz, Z, &iz, &fz);
    orc_pme_theta(order, fx, tx, d);
    orc_pme_theta(order, fy, ty, d);
    orc_pme_theta(order, fz, tz, d);
    const double q = xyzq[4 * i + 3];
    for (int c = 0; c < order; c++)
      for (int b = 0; b < order; b++)
        for (int a = 0; a < order; a++) {
          const size_t p = wrapp(ix + a, X) + ((size_t)wrapp(iy + b, Y) + (size_t)wrapp(iz + c, Z) * Y) * X;
          Q[p] += q * tx[a] * ty[b] * tz[c];
        }
  }
}

/* one scalar_sum term: eterm of spectrum index (k1, k2, k3) */
static double pme_eterm(int k1, int k2, int k3, int X, int Y, int Z, const double *r,
                        const double *p1, const double *p2, const double *p3, double fac,
                        double piv_inv, double *mh) {
  const int nf1 = X / 2 + X % 2, nf2 = Y / 2 + Y % 2, nf3 = Z / 2 + Z % 2;
  const int m1 = k1 >= nf1 ? k1 - X : k1, m2 = k2 >= nf2 ? k2 - Y : k2,
            m3 = k3 >= nf3 ? k3 - Z : k3;
  mh[0] = r[0] * m1;
  mh[1] = r[1] * m2;
  mh[2] = r[2] * m3;
  const double msq = mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2];
  return exp(-fac * msq) * piv_inv * p1[k1] * p2[k2] * p3[k3] / msq;
}

/* scalar_sum_ortho_kernel (:946-1256) on the r2c half spectrum Qk
 * ((X/2+1) x Y x Z complex, interleaved); energy/vir9 may be NULL */
void orc_pme_scalar_sum(int X, int Y, int Z, int order, double bx, double by, double bz,
                        double kappa, double ccelec, double *Qk, double *energy,
                        double *vir9) {
  double *p1 = (double *)malloc(sizeof(double) * X);
  double *p2 = (double *)malloc(sizeof(double) * Y);
  double *p3 = (double *)malloc(sizeof(double) * Z);
  orc_pme_prefac(order, X, p1);
  orc_pme_prefac(order, Y, p2);
  orc_pme_prefac(order, Z, p3);
  const double r[3] = {1.0 / bx, 1.0 / by, 1.0 / bz};
  const double piv_inv = r[0] * r[1] * r[2] / PI_S;
  const double fac = PI_S * PI_S / (kappa * kappa);
  const int size1 = X / 2 + 1;
  double E = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
  for (int k3 = 0; k3 < Z; k3++)
    for (int k2 = 0; k2 < Y; k2++)
      for (int k1 = 0; k1 < size1; k1++) {
        double *q = Qk + 2 * ((size_t)k1 + ((size_t)k2 + (size_t)k3 * Y) * size1);
        if (k1 == 0 && k2 == 0 && k3 == 0) {
          q[0] = 0.0;
          q[1] = 0.0;
          continue;
        }
        double mh[3];
        const double et = pme_eterm(k1, k2, k3, X, Y, Z, r, p1, p2, p3, fac, piv_inv, mh);
        const double q2 = q[0] * q[0] + q[1] * q[1];
        for (int part = 0; part < (k1 >= 1 ? 2 : 1); part++) {
          double e = et, m[3] = {mh[0], mh[1], mh[2]};
          if (part == 1) {
            const int k1s = X - k1, k2s = (Y - k2) % Y, k3s = (Z - k3) % Z;
            e = pme_eterm(k1s, k2s, k3s, X, Y, Z, r, p1, p2, p3, fac, piv_inv, m);
          }
          const double msq = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
          const double t1 = e * q2, vt = 2.0 * (fac + 1.0 / msq), t2 = t1 * vt;
          E += t1;
          v[0] += t1 * (vt * m[0] * m[0] - 1.0);
          v[1] += t2 * m[0] * m[1];
          v[2] += t2 * m[0] * m[2];
          v[3] += t1 * (vt * m[1] * m[1] - 1.0);
          v[4] += t2 * m[1] * m[2];
          v[5] += t1 * (vt * m[2] * m[2] - 1.0);
        }
        q[0] *= et;
        q[1] *= et;
      }
  const double h = 0.5 * ccelec;
  if (energy) *energy = E * h;
  if (vir9) {
    vir9[0] = -h * v[0];
    vir9[1] = -h * v[1];
    vir9[2] = -h * v[2];
    vir9[3] = -h * v[1];
    vir9[4] = -h * v[3];
    vir9[5] = -h * v[4];
    vir9[6] = -h * v[2];
    vir9[7] = -h * v[4];
    vir9[8] = -h * v[5];
  }
  free(p1);
  free(p2);
  free(p3);
}

/* gather_force_<order>_ortho_kernel (:1612-1890): the energy gradient
 * dE/dr (the reference's gather output, negated into the force by its
 * driver) from the convolved grid Qc */
void orc_pme_gather(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                    double by, double bz, double ccelec, const double *Qc, double *gx,
                    double *gy, double *gz) {
  double tx[16], ty[16], tz[16], dx[16], dy[16], dz[16];
  for (int i = 0; i < n; i++) {
    int ix, iy, iz;
    double fx, fy, fz;
    pme_pos(xyzq[4 * i + 0], 1.0 / bx, X, &ix, &fx);
    pme_pos(xyzq[4 * i + 1], 1.0 / by, Y, &iy, &fy);
    pme_pos(xyzq[4 * i + 2], 1.0 / bz, Z, &iz, &fz);
    orc_pme_theta(order, fx, tx, dx);
    orc_pme_theta(order, fy, ty, dy);
    orc_pme_theta(order, fz, tz, dz);
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int c = 0; c < order; c++)
      for (int b = 0; b < order; b++)
        for (int a = 0; a < order; a++) {
          const double g =
              Qc[wrapp(ix + a, X) + ((size_t)wrapp(iy + b, Y) + (size_t)wrapp(iz + c, Z) * Y) * X];
          f1 += dx[a] * ty[b] * tz[c] * g;
          f2 += tx[a] * dy[b] * tz[c] * g;
          f3 += tx[a] * ty[b] * dz[c] * g;
        }
    const double q = xyzq[4 * i + 3] * ccelec;
    gx[i] = q * (1.0 / bx) * f1 * X;
    gy[i] = q * (1.0 / by) * f2 * Y;
    gz[i] = q * (1.0 / bz) * f3 * Z;
  }
}

/* The comparator's step as gpu_gse.cu:61-73 drives it: spread -> r2c ->
 * scalar_sum (energy, virial) -> c2r -> gather; forces = -dE/dr.  Any output
 * may be NULL. */
void orc_pme_step(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                  double by, double bz, double kappa, double ccelec, double *energy,
                  double *vir9, double *fx, double *fy, double *fz, double *Q_out) {
  const size_t sz = (size_t)X * Y * Z;
  double *Q = (double *)malloc(sizeof(double) * sz);
  double *Qk = (double *)malloc(sizeof(double) * 2 * (size_t)(X / 2 + 1) * Y * Z);
  orc_pme_spread(X, Y, Z, order, n, xyzq, bx, by, bz, Q);
  if (Q_out) memcpy(Q_out, Q, sizeof(double) * sz);
  orc_fft_r2c(X, Y, Z, Q, Qk);
  orc_pme_scalar_sum(X, Y, Z, order, bx, by, bz, kappa, ccelec, Qk, energy, vir9);
  if (fx || fy || fz) {
    orc_fft_c2r(X, Y, Z, Qk, Q);
    double *g = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    orc_pme_gather(X, Y, Z, order, n, xyzq, bx, by, bz, ccelec, Q, g, g + n, g + 2 * n);
    for (int i = 0; i < n; i++) {
      if (fx) fx[i] = -g[i];
      if (fy) fy[i] = -g[n + i];
      if (fz) fz[i] = -g[2 * n + i];
    }
    free(g);
  }
  free(Q);
  free(Qk);
}
