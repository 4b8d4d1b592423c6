/*
 * OpenMP timing build of the oracle's GSE-k step -- TEST INFRASTRUCTURE,
 * used only by bench.py's cpu_baseline leg to time the CPU path across all
 * host cores.  It is the same algorithm as orc_gse_step (gse_oracle.c, which
 * this file includes) with the work split over threads:
 *   spread   thread-private grids summed afterwards (CpuGaussCharge.cpp:235-384)
 *   FFTs     1D lines of each axis in parallel (CpuFFTSolver.cpp:75-95)
 *   G(k)     the half spectrum in parallel (CpuGreensFuncGSE.cpp:9-56)
 *   energy   parallel sum of rho*phi (CpuGSE.cpp:86-104)
 *   gather   particles in parallel: the sphere force of
 *            CpuGaussCharge.cpp:105-172 and the cube potential phi_i
 * Sums run in a different order than the serial oracle, so results agree
 * with it to rounding, not bit for bit; the parity checks use the serial
 * oracle only.
 */
#include <omp.h>

#include "gse_oracle.c"

static void fft_axis_omp(int X, int Y, int Z, cpx *c, int inv, int axis) {
  const int n = axis == 0 ? X : (axis == 1 ? Y : Z);
  const int lines = axis == 0 ? Y * Z : (axis == 1 ? X * Z : X * Y);
  plan1d pl;
  plan_make(&pl, n);
  const int mf = max_factor(&pl);
#pragma omp parallel
  {
    cpx *line = (cpx *)malloc(sizeof(cpx) * n);
    cpx *work = (cpx *)malloc(sizeof(cpx) * n);
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * 2 * mf);
#pragma omp for schedule(static)
    for (int l = 0; l < lines; l++) {
      size_t base, stride;
      if (axis == 0) {
        base = (size_t)l * X;
        stride = 1;
      } else if (axis == 1) {
        const int x = l % X, z = l / X;
        base = x + (size_t)z * Y * X;
        stride = X;
      } else {
        base = (size_t)l;
        stride = (size_t)X * Y;
      }
      for (int j = 0; j < n; j++) line[j] = c[base + j * stride];
      fft_run(&pl, line, inv, work, tmp);
      for (int j = 0; j < n; j++) c[base + j * stride] = line[j];
    }
    free(line);
    free(work);
    free(tmp);
  }
  plan_free(&pl);
}

static void fft3d_omp(int X, int Y, int Z, cpx *c, int inv) {
  fft_axis_omp(X, Y, Z, c, inv, 0);
  fft_axis_omp(X, Y, Z, c, inv, 1);
  fft_axis_omp(X, Y, Z, c, inv, 2);
}

/* One GSE-k step (spread sigma1 -> r2c -> G_GSE(sigma2) -> c2r -> energy ->
 * sphere force x ccelec and phi_i x ccelec).  Returns the energy. */
double orc_step_omp(const orc_gse_params *p, int n, const double *xyzq, double *fx,
                    double *fy, double *fz, double *pot) {
  const int X = p->X, Y = p->Y, Z = p->Z;
  const size_t sz = (size_t)X * Y * Z;
  double s1, s2, rc1, rc2;
  orc_gse_derive(p, &s1, &s2, &rc1, &rc2);
  const double bx = p->bx, by = p->by, bz = p->bz;
  const double hx = bx / X, hy = by / Y, hz = bz / Z;
  const int nx = (int)ceil(rc1 / hx), ny = (int)ceil(rc1 / hy), nz = (int)ceil(rc1 / hz);
  const double inv_2s2 = 1.0 / (2.0 * s1 * s1);
  const double pref = pow(2.0 * PI_D * s1 * s1, -1.5);
  if (2 * nx + 1 > 64 || 2 * ny + 1 > 64 || 2 * nz + 1 > 64) return NAN;
  const int nt = omp_get_max_threads();
  double *rho = (double *)calloc(sz, sizeof(double));
  double *priv = (double *)malloc(sizeof(double) * sz * (size_t)nt);
  /* spread: each thread accumulates its share of particles in its own grid */
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    double *g = priv + (size_t)t * sz;
    memset(g, 0, sizeof(double) * sz);
    double wx[64], wy[64], wz[64];
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) {
      const double x = xyzq[4 * i], y = xyzq[4 * i + 1], z = xyzq[4 * i + 2];
      const double qp = xyzq[4 * i + 3] * pref;
      int ixc = (int)round(x / hx) - nx, iyc = (int)round(y / hy) - ny,
          izc = (int)round(z / hz) - nz;
      const double dx = ixc * hx - x, dy = iyc * hy - y, dz = izc * hz - z;
      for (int j = 0; j <= 2 * nx; j++) wx[j] = exp(-(dx + j * hx) * (dx + j * hx) * inv_2s2);
      for (int j = 0; j <= 2 * ny; j++) wy[j] = exp(-(dy + j * hy) * (dy + j * hy) * inv_2s2);
      for (int j = 0; j <= 2 * nz; j++) wz[j] = exp(-(dz + j * hz) * (dz + j * hz) * inv_2s2);
      ixc = wrapi(ixc, X);
      iyc = wrapi(iyc, Y);
      izc = wrapi(izc, Z);
      for (int tz = 0; tz <= 2 * nz; tz++)
        for (int ty = 0; ty <= 2 * ny; ty++) {
          double *row = g + ((size_t)((iyc + ty) % Y) + (size_t)((izc + tz) % Z) * Y) * X;
          const double qyz = qp * wy[ty] * wz[tz];
          for (int tx = 0; tx <= 2 * nx; tx++) row[(ixc + tx) % X] += qyz * wx[tx];
        }
    }
  }
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < sz; k++) {
    double s = 0.0;
    for (int t = 0; t < nt; t++) s += priv[(size_t)t * sz + k];
    rho[k] = s;
  }
  free(priv);
  /* r2c (complex transform of the real grid), G multiply, c2r */
  cpx *c = (cpx *)malloc(sizeof(cpx) * sz);
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < sz; k++) {
    c[k].re = rho[k];
    c[k].im = 0.0;
  }
  fft3d_omp(X, Y, Z, c, 0);
  const int hX = X / 2 + X % 2, hY = Y / 2 + Y % 2, hZ = Z / 2 + Z % 2;
  const double fourpi_vol = 4.0 * PI_D / (double)(X * Y * Z);
  const double sf = s2 * s2 / 2.0;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < Z; iz++)
    for (int iy = 0; iy < Y; iy++)
      for (int ix = 0; ix < X; ix++) {
        const int kx = ix >= hX ? ix - X : ix, ky = iy >= hY ? iy - Y : iy,
                  kz = iz >= hZ ? iz - Z : iz;
        const double k1 = kx * 2.0 * PI_D / bx, k2v = ky * 2.0 * PI_D / by,
                     k3 = kz * 2.0 * PI_D / bz;
        const double k2 = k1 * k1 + k2v * k2v + k3 * k3;
        const double g = (ix == 0 && iy == 0 && iz == 0) ? 0.0 : fourpi_vol * exp(-sf * k2) / k2;
        cpx *v = &c[ix + (size_t)(iy + (size_t)iz * Y) * X];
        v->re *= g;
        v->im *= g;
      }
  fft3d_omp(X, Y, Z, c, 1);
  double *phi = (double *)malloc(sizeof(double) * sz);
  double E = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : E)
  for (size_t k = 0; k < sz; k++) {
    phi[k] = c[k].re;
    E += rho[k] * phi[k];
  }
  E *= p->ccelec * hx * hy * hz * 0.5;
  free(c);
  /* gather: sphere force and cube potential per particle */
  const double prefF = pow(2.0 * PI_D, -1.5) * hx * hy * hz * pow(s1, -5.0);
  const double prefP = pref * hx * hy * hz;
  const double rcut2 = rc1 * rc1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    const double x = xyzq[4 * i], y = xyzq[4 * i + 1], z = xyzq[4 * i + 2];
    const double q = xyzq[4 * i + 3];
    const int ix = (int)round(x / hx), iy = (int)round(y / hy), iz = (int)round(z / hz);
    double rx = 0.0, ry = 0.0, rz = 0.0, ps = 0.0;
    for (int tz = iz - nz; tz <= iz + nz; tz++) {
      const double dz = z - tz * hz;
      const int pz = wrapi(tz, Z);
      for (int ty = iy - ny; ty <= iy + ny; ty++) {
        const double dy = y - ty * hy;
        const double *row = phi + ((size_t)wrapi(ty, Y) + (size_t)pz * Y) * X;
        for (int tx = ix - nx; tx <= ix + nx; tx++) {
          const double dx = x - tx * hx;
          const double r2 = dx * dx + dy * dy + dz * dz;
          const double w = exp(-r2 * inv_2s2) * row[wrapi(tx, X)];
          ps += w;
          if (r2 < rcut2) {
            rx += w * dx;
            ry += w * dy;
            rz += w * dz;
          }
        }
      }
    }
    const double fs = q * prefF * p->ccelec;
    fx[i] = rx * fs;
    fy[i] = ry * fs;
    fz[i] = rz * fs;
    pot[i] = ps * prefP * p->ccelec;
  }
  free(rho);
  free(phi);
  return E;
}

int orc_omp_threads(void) { return omp_get_max_threads(); }

void orc_omp_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
