/*
 * FP64 CPU oracle for the GSE reciprocal-space path -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off so that every product
 * and sum is rounded exactly where the reference's C++ rounds it).
 *
 * Each function restates one routine of the reference (file:line given at
 * the definition).  Loop order, operand order and the index arithmetic are
 * kept, because the cube selection ((int)round(x*inv_h)) and the sphere test
 * (r2 < rcut2) are discrete decisions that must come out the same.
 */
#include "gse_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI_D = 3.14159265358979323846; /* TypesGSE.h pi_dbl */

static int wrapi(int i, int m) { /* non-negative modulo */
  int r = i % m;
  return r < 0 ? r + m : r;
}

/* ------------------------------------------------------------------------ */
/* CpuGaussCharge.cpp:28-56 calcSupportSize                                  */
/* ------------------------------------------------------------------------ */
int orc_support_size(int X, int Y, int Z, double rcut, double bx, double by,
                     double bz) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const int nx = (int)ceil(rcut / hxd), ny = (int)ceil(rcut / hyd),
            nz = (int)ceil(rcut / hzd);
  const double rcut2 = rcut * rcut;
  int nsup = 0;
  for (int tz = -nz; tz <= nz; tz++) {
    double dz = ((double)(abs(tz) - 1)) * hzd;
    for (int ty = -ny; ty <= ny; ty++) {
      double dy = ((double)(abs(ty) - 1)) * hyd;
      for (int tx = -nx; tx <= nx; tx++) {
        double dx = ((double)(abs(tx) - 1)) * hxd;
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < rcut2) nsup++;
      }
    }
  }
  return nsup;
}

/* ------------------------------------------------------------------------ */
/* CpuGaussCharge.cpp:235-384 spreadChargeToGrid                             */
/* ------------------------------------------------------------------------ */
void orc_spread(int X, int Y, int Z, double sigma, double rcut, int n,
                const double *xyzq, double bx, double by, double bz,
                double *rho, int normalize) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const int nx = (int)ceil(rcut / hxd), ny = (int)ceil(rcut / hyd),
            nz = (int)ceil(rcut / hzd);
  const double inv_hx = 1.0 / hxd, inv_hy = 1.0 / hyd, inv_hz = 1.0 / hzd;
  const double hx = hxd, hy = hyd, hz = hzd;
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double pref = pow(2.0 * PI_D * sigma * sigma, -3.0 / 2.0);

  memset(rho, 0, sizeof(double) * (size_t)X * Y * Z);
  double *wx = (double *)malloc(sizeof(double) * (2 * nx + 1));
  double *wy = (double *)malloc(sizeof(double) * (2 * ny + 1));
  double *wz = (double *)malloc(sizeof(double) * (2 * nz + 1));

  for (int i = 0; i < n; i++) {
    const double x = xyzq[4 * i + 0], y = xyzq[4 * i + 1], z = xyzq[4 * i + 2];
    const double q_pref = xyzq[4 * i + 3] * pref;
    int ixc = (int)round(x * inv_hx) - nx;
    int iyc = (int)round(y * inv_hy) - ny;
    int izc = (int)round(z * inv_hz) - nz;
    const double dx = ixc * hx - x;
    const double dy = iyc * hy - y;
    const double dz = izc * hz - z;
    for (int j = 0; j <= 2 * nx; j++) {
      double t = dx + j * hx;
      wx[j] = exp(-t * t * inv_2sigmasq);
    }
    for (int j = 0; j <= 2 * ny; j++) {
      double t = dy + j * hy;
      wy[j] = exp(-t * t * inv_2sigmasq);
    }
    for (int j = 0; j <= 2 * nz; j++) {
      double t = dz + j * hz;
      wz[j] = exp(-t * t * inv_2sigmasq);
    }
    if (normalize) { /* :330-362, shift is zero in the reference */
      const double pref1 = sqrt(2.0 * PI_D * sigma * sigma);
      double nrm = 0.0;
      for (int j = 0; j <= 2 * nx; j++) nrm += wx[j];
      nrm = pref1 / nrm;
      for (int j = 0; j <= 2 * nx; j++) wx[j] *= nrm;
      nrm = 0.0;
      for (int j = 0; j <= 2 * ny; j++) nrm += wy[j];
      nrm = pref1 / nrm;
      for (int j = 0; j <= 2 * ny; j++) wy[j] *= nrm;
      nrm = 0.0;
      for (int j = 0; j <= 2 * nz; j++) nrm += wz[j];
      nrm = pref1 / nrm;
      for (int j = 0; j <= 2 * nz; j++) wz[j] *= nrm;
    }
    /* :364-366 wraps with (c+M)%M; a full modulo is identical wherever the
     * reference index is non-negative and stays defined elsewhere */
    ixc = wrapi(ixc, X);
    iyc = wrapi(iyc, Y);
    izc = wrapi(izc, Z);
    for (int tz = 0; tz <= 2 * nz; tz++) {
      for (int ty = 0; ty <= 2 * ny; ty++) {
        for (int tx = 0; tx <= 2 * nx; tx++) {
          int ix = (ixc + tx) % X;
          int iy = (iyc + ty) % Y;
          int iz = (izc + tz) % Z;
          const int p = ix + (iy + iz * Y) * X;
          rho[p] += q_pref * wx[tx] * wy[ty] * wz[tz];
        }
      }
    }
  }
  free(wx);
  free(wy);
  free(wz);
}

/* ------------------------------------------------------------------------ */
/* separable periodic convolution passes (gather form)                      */
/* ------------------------------------------------------------------------ */
static void conv_pass_x(int X, int Y, int Z, int R, const double *k,
                        const double *in, double *out) {
  for (int pz = 0; pz < Z; pz++)
    for (int py = 0; py < Y; py++) {
      int pyz = (py + pz * Y) * X;
      for (int ix = 0; ix < X; ix++) {
        double sum = 0.0;
        for (int ii = -R; ii <= R; ii++) {
          int px = wrapi(ii + ix, X);
          sum += k[ii + R] * in[px + pyz];
        }
        out[ix + pyz] = sum;
      }
    }
}

static void conv_pass_y(int X, int Y, int Z, int R, const double *k,
                        const double *in, double *out) {
  for (int pz = 0; pz < Z; pz++)
    for (int px = 0; px < X; px++)
      for (int iy = 0; iy < Y; iy++) {
        double sum = 0.0;
        for (int ii = -R; ii <= R; ii++) {
          int py = wrapi(ii + iy, Y);
          sum += k[ii + R] * in[px + (py + pz * Y) * X];
        }
        out[px + (iy + pz * Y) * X] = sum;
      }
}

static void conv_pass_z(int X, int Y, int Z, int R, const double *k,
                        const double *in, double *out) {
  for (int px = 0; px < X; px++)
    for (int py = 0; py < Y; py++)
      for (int iz = 0; iz < Z; iz++) {
        double sum = 0.0;
        for (int ii = -R; ii <= R; ii++) {
          int pz = wrapi(ii + iz, Z);
          sum += k[ii + R] * in[px + (py + pz * Y) * X];
        }
        out[px + (py + iz * Y) * X] = sum;
      }
}

/* CpuGaussCharge.cpp:390-519 spreadChargeOnGrid */
void orc_spread_on_grid(int X, int Y, int Z, double sigma, double rcut,
                        double bx, double by, double bz, const double *rho,
                        double *rhoS) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const int nx = (int)ceil(rcut / hxd), ny = (int)ceil(rcut / hyd),
            nz = (int)ceil(rcut / hzd);
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double pref = pow(2.0 * PI_D * sigma * sigma, -3.0 / 2.0) * hxd * hyd * hzd;
  const double hxsq = hxd * hxd, hysq = hyd * hyd, hzsq = hzd * hzd;
  double *gx = (double *)malloc(sizeof(double) * (2 * nx + 1));
  double *gy = (double *)malloc(sizeof(double) * (2 * ny + 1));
  double *gz = (double *)malloc(sizeof(double) * (2 * nz + 1));
  for (int ii = -nx; ii <= nx; ii++) gx[ii + nx] = exp(-hxsq * (double)(ii * ii) * inv_2sigmasq);
  for (int ii = -ny; ii <= ny; ii++) gy[ii + ny] = exp(-hysq * (double)(ii * ii) * inv_2sigmasq);
  for (int ii = -nz; ii <= nz; ii++) gz[ii + nz] = pref * exp(-hzsq * (double)(ii * ii) * inv_2sigmasq);
  size_t sz = (size_t)X * Y * Z;
  double *tx = (double *)malloc(sizeof(double) * sz);
  double *txy = (double *)malloc(sizeof(double) * sz);
  conv_pass_x(X, Y, Z, nx, gx, rho, tx);
  conv_pass_y(X, Y, Z, ny, gy, tx, txy);
  conv_pass_z(X, Y, Z, nz, gz, txy, rhoS);
  free(tx);
  free(txy);
  free(gx);
  free(gy);
  free(gz);
}

/* CpuConvolution.cpp:116-151: X pass, then dataXYZ.copy(dataX) */
void orc_conv_run_x(int X, int Y, int Z, int Rx, const double *kx,
                    const double *in, double *out) {
  conv_pass_x(X, Y, Z, Rx, kx, in, out);
}

void orc_conv_run_xyz(int X, int Y, int Z, int Rx, int Ry, int Rz,
                      const double *kx, const double *ky, const double *kz,
                      const double *in, double *out) {
  size_t sz = (size_t)X * Y * Z;
  double *tx = (double *)malloc(sizeof(double) * sz);
  double *txy = (double *)malloc(sizeof(double) * sz);
  conv_pass_x(X, Y, Z, Rx, kx, in, tx);
  conv_pass_y(X, Y, Z, Ry, ky, tx, txy);
  conv_pass_z(X, Y, Z, Rz, kz, txy, out);
  free(tx);
  free(txy);
}

/* ------------------------------------------------------------------------ */
/* CpuGaussCharge.cpp:105-172 interpolateLoop (sphere support)              */
/* ------------------------------------------------------------------------ */
void orc_interpolate(int X, int Y, int Z, double sigma, double rcut, int n,
                     const double *xyzq, double bx, double by, double bz,
                     const double *phi, double *resX, double *resY,
                     double *resZ) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const int nx = (int)ceil(rcut / hxd), ny = (int)ceil(rcut / hyd),
            nz = (int)ceil(rcut / hzd);
  const double inv_hx = 1.0 / hxd, inv_hy = 1.0 / hyd, inv_hz = 1.0 / hzd;
  const double hx = hxd, hy = hyd, hz = hzd;
  const double rcut2 = rcut * rcut;
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double pref = pow(2.0 * PI_D, -3.0 / 2.0) * hxd * hyd * hzd * pow(sigma, -5.0);
  for (int i = 0; i < n; i++) {
    const double x = xyzq[4 * i + 0], y = xyzq[4 * i + 1], z = xyzq[4 * i + 2];
    const double fac = xyzq[4 * i + 3] * pref;
    int ix = (int)round(x * inv_hx);
    int iy = (int)round(y * inv_hy);
    int iz = (int)round(z * inv_hz);
    double rx = 0.0, ry = 0.0, rz = 0.0;
    for (int tz = iz - nz; tz <= iz + nz; tz++) {
      double zp = ((double)tz) * hz;
      double dz = z - zp;
      int pz = wrapi(tz, Z);
      for (int ty = iy - ny; ty <= iy + ny; ty++) {
        double yp = ((double)ty) * hy;
        double dy = y - yp;
        int py = wrapi(ty, Y);
        for (int tx = ix - nx; tx <= ix + nx; tx++) {
          double xp = ((double)tx) * hx;
          double dx = x - xp;
          int px = wrapi(tx, X);
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < rcut2) {
            int p = px + (py + pz * Y) * X;
            double r = fac * exp(-r2 * inv_2sigmasq) * phi[p];
            rx += r * dx;
            ry += r * dy;
            rz += r * dz;
          }
        }
      }
    }
    resX[i] = rx;
    resY[i] = ry;
    resZ[i] = rz;
  }
}

/* Build-defined cube gather: the exact adjoint of orc_spread (normalize=0).
 * s_ip = q_i*pref*h^3/sigma^2 * w_ip * phi_p with w_ip the Gaussian weight
 * over the spread's (2n+1)^3 cube; F_i = sum_p s_ip d_ip, d = r_i - r_p;
 * phi_i = pref*h^3 * sum_p w_ip phi_p; vir_ab = sum_i sum_p s_ip d_a d_b. */
void orc_gather_cube(int X, int Y, int Z, double sigma, double rcut, int n,
                     const double *xyzq, double bx, double by, double bz,
                     const double *phi, double *pot, double *fx, double *fy,
                     double *fz, double *vir9) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const int nx = (int)ceil(rcut / hxd), ny = (int)ceil(rcut / hyd),
            nz = (int)ceil(rcut / hzd);
  const double inv_hx = 1.0 / hxd, inv_hy = 1.0 / hyd, inv_hz = 1.0 / hzd;
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double prefP = pow(2.0 * PI_D * sigma * sigma, -3.0 / 2.0) * hxd * hyd * hzd;
  const double prefF = prefP / (sigma * sigma);
  double V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    const double x = xyzq[4 * i + 0], y = xyzq[4 * i + 1], z = xyzq[4 * i + 2];
    const double q = xyzq[4 * i + 3];
    int ix = (int)round(x * inv_hx);
    int iy = (int)round(y * inv_hy);
    int iz = (int)round(z * inv_hz);
    double sp = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int tz = iz - nz; tz <= iz + nz; tz++) {
      double dz = z - ((double)tz) * hzd;
      int pz = wrapi(tz, Z);
      for (int ty = iy - ny; ty <= iy + ny; ty++) {
        double dy = y - ((double)ty) * hyd;
        int py = wrapi(ty, Y);
        for (int tx = ix - nx; tx <= ix + nx; tx++) {
          double dx = x - ((double)tx) * hxd;
          int px = wrapi(tx, X);
          double r2 = dx * dx + dy * dy + dz * dz;
          double wphi = exp(-r2 * inv_2sigmasq) * phi[px + (py + pz * Y) * X];
          sp += wphi;
          rx += wphi * dx;
          ry += wphi * dy;
          rz += wphi * dz;
          if (vir9) {
            double d[3] = {dx, dy, dz};
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) v[3 * a + b] += wphi * d[a] * d[b];
          }
        }
      }
    }
    if (pot) pot[i] = prefP * sp;
    if (fx) fx[i] = q * prefF * rx;
    if (fy) fy[i] = q * prefF * ry;
    if (fz) fz[i] = q * prefF * rz;
    if (vir9)
      for (int k = 0; k < 9; k++) V[k] += q * prefF * v[k];
  }
  if (vir9)
    for (int k = 0; k < 9; k++) vir9[k] = V[k];
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, twiddles from one table   */
/* ------------------------------------------------------------------------ */
typedef struct {
  double re, im;
} cpx;

typedef struct {
  int n;
  int nf;
  int f[64];
  cpx *tw; /* tw[k] = exp(-2 pi i k / n) */
} plan1d;

static void plan_make(plan1d *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  for (int r = 3; m > 1; r += 2)
    while (m % r == 0) { p->f[p->nf++] = r; m /= r; }
  if (p->nf == 0) p->f[p->nf++] = 1;
  p->tw = (cpx *)malloc(sizeof(cpx) * n);
  for (int k = 0; k < n; k++) {
    double a = 2.0 * PI_D * (double)k / (double)n;
    p->tw[k].re = cos(a);
    p->tw[k].im = -sin(a);
  }
}

static void plan_free(plan1d *p) { free(p->tw); }

/* out[k] = sum_j in[j*is] W_n^{jk}; W_n^x = tw[(x mod n) * ts] (conj if inv) */
static void fft_rec(const cpx *in, int is, cpx *out, int n, const int *f,
                    const cpx *tw, int ts, int N, int inv, cpx *tmp) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = f[0], m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + (size_t)q * is, is * p, out + (size_t)q * m, m, f + 1, tw,
            ts * p, N, inv, tmp);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) { /* t_q = W_n^{qk} Y_q[k] */
      cpx y = out[q * m + k];
      int e = (int)(((long)q * k) % n) * ts;
      cpx w = tw[e];
      if (inv) w.im = -w.im;
      tmp[q].re = y.re * w.re - y.im * w.im;
      tmp[q].im = y.re * w.im + y.im * w.re;
    }
    for (int s = 0; s < p; s++) { /* X[k+s m] = sum_q t_q W_p^{qs} */
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < p; q++) {
        int e = (int)(((long)q * s * m) % n) * ts;
        cpx w = tw[e];
        if (inv) w.im = -w.im;
        sr += tmp[q].re * w.re - tmp[q].im * w.im;
        si += tmp[q].re * w.im + tmp[q].im * w.re;
      }
      tmp[p + s].re = sr;
      tmp[p + s].im = si;
    }
    for (int s = 0; s < p; s++) out[k + s * m] = tmp[p + s];
  }
}

static void fft_run(const plan1d *pl, cpx *data, int inv, cpx *work, cpx *tmp) {
  memcpy(work, data, sizeof(cpx) * pl->n);
  fft_rec(work, 1, data, pl->n, pl->f, pl->tw, 1, pl->n, inv, tmp);
}

static int max_factor(const plan1d *p) {
  int m = 1;
  for (int i = 0; i < p->nf; i++)
    if (p->f[i] > m) m = p->f[i];
  return m;
}

void orc_fft_1d(int n, double *data, int sign) {
  plan1d p;
  plan_make(&p, n);
  cpx *work = (cpx *)malloc(sizeof(cpx) * n);
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * 2 * max_factor(&p));
  fft_run(&p, (cpx *)data, sign > 0, work, tmp);
  free(tmp);
  free(work);
  plan_free(&p);
}

/* full complex 3D transform of c[z][y][x] */
static void fft3d(int X, int Y, int Z, cpx *c, int inv) {
  plan1d px, py, pz;
  plan_make(&px, X);
  plan_make(&py, Y);
  plan_make(&pz, Z);
  int mx = X > Y ? X : Y;
  if (Z > mx) mx = Z;
  int mf = max_factor(&px);
  if (max_factor(&py) > mf) mf = max_factor(&py);
  if (max_factor(&pz) > mf) mf = max_factor(&pz);
  cpx *line = (cpx *)malloc(sizeof(cpx) * mx);
  cpx *work = (cpx *)malloc(sizeof(cpx) * mx);
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * 2 * mf);
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++) fft_run(&px, c + (size_t)(y + z * Y) * X, inv, work, tmp);
  for (int z = 0; z < Z; z++)
    for (int x = 0; x < X; x++) {
      for (int y = 0; y < Y; y++) line[y] = c[x + (size_t)(y + z * Y) * X];
      fft_run(&py, line, inv, work, tmp);
      for (int y = 0; y < Y; y++) c[x + (size_t)(y + z * Y) * X] = line[y];
    }
  for (int y = 0; y < Y; y++)
    for (int x = 0; x < X; x++) {
      for (int z = 0; z < Z; z++) line[z] = c[x + (size_t)(y + z * Y) * X];
      fft_run(&pz, line, inv, work, tmp);
      for (int z = 0; z < Z; z++) c[x + (size_t)(y + z * Y) * X] = line[z];
    }
  free(line);
  free(work);
  free(tmp);
  plan_free(&px);
  plan_free(&py);
  plan_free(&pz);
}

void orc_fft_r2c(int X, int Y, int Z, const double *in, double *outc) {
  size_t sz = (size_t)X * Y * Z;
  cpx *c = (cpx *)malloc(sizeof(cpx) * sz);
  for (size_t i = 0; i < sz; i++) {
    c[i].re = in[i];
    c[i].im = 0.0;
  }
  fft3d(X, Y, Z, c, 0);
  const int Xh = X / 2 + 1;
  cpx *o = (cpx *)outc;
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < Xh; x++) o[x + (size_t)(y + z * Y) * Xh] = c[x + (size_t)(y + z * Y) * X];
  free(c);
}

/* The half spectrum is completed by Hermitian mirroring and the real part of
 * the inverse transform is kept; for x = 0 and x = X/2 this equals FFTW's
 * treatment (imaginary parts of the 1D c2r end points are ignored). */
void orc_fft_c2r(int X, int Y, int Z, const double *inc, double *out) {
  size_t sz = (size_t)X * Y * Z;
  const int Xh = X / 2 + 1;
  const cpx *h = (const cpx *)inc;
  cpx *c = (cpx *)malloc(sizeof(cpx) * sz);
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < X; x++) {
        cpx v;
        if (x < Xh) {
          v = h[x + (size_t)(y + z * Y) * Xh];
        } else {
          int zz = wrapi(-z, Z), yy = wrapi(-y, Y), xx = X - x;
          v = h[xx + (size_t)(yy + zz * Y) * Xh];
          v.im = -v.im;
        }
        c[x + (size_t)(y + z * Y) * X] = v;
      }
  fft3d(X, Y, Z, c, 1);
  for (size_t i = 0; i < sz; i++) out[i] = c[i].re;
  free(c);
}

/* ------------------------------------------------------------------------ */
/* Green's functions: CpuGreensFuncGSE.cpp:9-56, CpuGreensFuncG2.cpp:9-60    */
/* ------------------------------------------------------------------------ */
void orc_greens_mul(int kind, int nfftX, int nfftY, int nfftZ, double bx,
                    double by, double bz, double sigma, double *rhoData) {
  const int sizeX = nfftX / 2 + 1, sizeY = nfftY, sizeZ = nfftZ;
  const int halfX = nfftX / 2 + (nfftX % 2);
  const int halfY = nfftY / 2 + (nfftY % 2);
  const int halfZ = nfftZ / 2 + (nfftZ % 2);
  const double twopi_bx = 2.0 * PI_D / bx, twopi_by = 2.0 * PI_D / by,
               twopi_bz = 2.0 * PI_D / bz;
  const double fourpi_vol = 4.0 * PI_D / (double)(nfftX * nfftY * nfftZ);
  const double inv_vol = 1.0 / (double)(nfftX * nfftY * nfftZ);
  const double sigmafac = sigma * sigma / 2.0;
  int p = 0;
  for (int iz = 0; iz < sizeZ; iz++)
    for (int iy = 0; iy < sizeY; iy++)
      for (int ix = 0; ix < sizeX; ix++, p += 2) {
        if (kind == 0 && p == 0) {
          rhoData[0] = 0.0;
          rhoData[1] = 0.0;
          continue;
        }
        int kx = (ix >= halfX) ? ix - nfftX : ix;
        int ky = (iy >= halfY) ? iy - nfftY : iy;
        int kz = (iz >= halfZ) ? iz - nfftZ : iz;
        double kxd = ((double)kx) * twopi_bx;
        double kyd = ((double)ky) * twopi_by;
        double kzd = ((double)kz) * twopi_bz;
        double kd2 = kxd * kxd + kyd * kyd + kzd * kzd;
        double fac = (kind == 0) ? fourpi_vol * exp(-sigmafac * kd2) / kd2
                                 : inv_vol * exp(-sigmafac * kd2);
        rhoData[p] *= fac;
        rhoData[p + 1] *= fac;
      }
}

/* CpuFFTSolver.cpp:75-95 */
void orc_fft_solve(int kind, int X, int Y, int Z, double sigma, double bx,
                   double by, double bz, const double *rho, double *phi) {
  double *rk = (double *)malloc(sizeof(double) * 2 * (size_t)(X / 2 + 1) * Y * Z);
  orc_fft_r2c(X, Y, Z, rho, rk);
  orc_greens_mul(kind, X, Y, Z, bx, by, bz, sigma, rk);
  orc_fft_c2r(X, Y, Z, rk, phi);
  free(rk);
}

/* CpuGSE.cpp:86-104 */
double orc_energy(int X, int Y, int Z, double ccelec, double bx, double by,
                  double bz, const double *rho, const double *phi) {
  const int size = X * Y * Z;
  double energy = 0.0;
  for (int i = 0; i < size; i++) energy += rho[i] * phi[i];
  const double hx = bx / (double)X, hy = by / (double)Y, hz = bz / (double)Z;
  energy *= ccelec * hx * hy * hz * 0.5;
  return energy;
}

/* Parseval form of CpuGSE.cpp:86-104 over the half spectrum, with the
 * k-space virial -(d E / d eps_ab) restricted to the G(k) dependence and
 * the cell-volume factor:  W^k_ab = -delta_ab E - sum_k E_k 2(s+1/k^2) k_a k_b
 * (s = sigma^2/2; the 1/k^2 term is absent for kind 1). */
double orc_energy_kspace(int kind, int nfftX, int nfftY, int nfftZ,
                         double ccelec, double bx, double by, double bz,
                         double sigma, const double *rhoK, double *vir9) {
  const int sizeX = nfftX / 2 + 1;
  const int halfX = nfftX / 2 + (nfftX % 2);
  const int halfY = nfftY / 2 + (nfftY % 2);
  const int halfZ = nfftZ / 2 + (nfftZ % 2);
  const double twopi_bx = 2.0 * PI_D / bx, twopi_by = 2.0 * PI_D / by,
               twopi_bz = 2.0 * PI_D / bz;
  const double fourpi_vol = 4.0 * PI_D / (double)(nfftX * nfftY * nfftZ);
  const double inv_vol = 1.0 / (double)(nfftX * nfftY * nfftZ);
  const double sigmafac = sigma * sigma / 2.0;
  const double vcell = (bx / nfftX) * (by / nfftY) * (bz / nfftZ);
  double E = 0.0;
  double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  size_t p = 0;
  for (int iz = 0; iz < nfftZ; iz++)
    for (int iy = 0; iy < nfftY; iy++)
      for (int ix = 0; ix < sizeX; ix++, p += 2) {
        if (kind == 0 && p == 0) continue;
        int kx = (ix >= halfX) ? ix - nfftX : ix;
        int ky = (iy >= halfY) ? iy - nfftY : iy;
        int kz = (iz >= halfZ) ? iz - nfftZ : iz;
        double kv[3] = {kx * twopi_bx, ky * twopi_by, kz * twopi_bz};
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double g = (kind == 0) ? fourpi_vol * exp(-sigmafac * k2) / k2
                               : inv_vol * exp(-sigmafac * k2);
        double mult = (ix == 0 || 2 * ix == nfftX) ? 1.0 : 2.0;
        double ek = 0.5 * ccelec * vcell * mult * g *
                    (rhoK[p] * rhoK[p] + rhoK[p + 1] * rhoK[p + 1]);
        E += ek;
        if (vir9) {
          double c = 2.0 * (sigmafac + (kind == 0 ? 1.0 / k2 : 0.0));
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) K[3 * a + b] += ek * c * kv[a] * kv[b];
        }
      }
  if (vir9)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) vir9[3 * a + b] = -(a == b ? E : 0.0) - K[3 * a + b];
  return E;
}

/* ------------------------------------------------------------------------ */
/* CpuGaussCharge.cpp:62-99 calcElectricFieldOnGrid                          */
/* ------------------------------------------------------------------------ */
void orc_efield(int X, int Y, int Z, double bx, double by, double bz,
                const double *phi, double *ex, double *ey, double *ez) {
  const double hxd = bx / (double)X, hyd = by / (double)Y, hzd = bz / (double)Z;
  const double inv_hx = 1.0 / hxd, inv_hy = 1.0 / hyd, inv_hz = 1.0 / hzd;
#define PHI(a, b, c) phi[wrapi(a, X) + (wrapi(b, Y) + wrapi(c, Z) * Y) * X]
  for (int iz = 0; iz < Z; iz++)
    for (int iy = 0; iy < Y; iy++)
      for (int ix = 0; ix < X; ix++) {
        int p = ix + (iy + iz * Y) * X;
        ex[p] = -(PHI(ix + 1, iy, iz) - PHI(ix, iy, iz)) * inv_hx;
        ey[p] = -(PHI(ix, iy + 1, iz) - PHI(ix, iy, iz)) * inv_hy;
        ez[p] = -(PHI(ix, iy, iz + 1) - PHI(ix, iy, iz)) * inv_hz;
      }
#undef PHI
}

/* CpuGaussCharge.cpp:189-226 calcDipoleSum */
void orc_dipole_sum(int X, int Y, int Z, double bx, double by, double bz,
                    const double *rho, double *dip3) {
  const double hx = bx / (double)X, hy = by / (double)Y, hz = bz / (double)Z;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int iz = 0; iz < Z; iz++) {
    double z = hz * (double)iz;
    for (int iy = 0; iy < Y; iy++) {
      double y = hy * (double)iy;
      for (int ix = 0; ix < X; ix++) {
        double x = hx * (double)ix;
        int p = ix + (iy + iz * Y) * X;
        double val = rho[p] * hx * hy * hz;
        sx += val * x;
        sy += val * y;
        sz += val * z;
      }
    }
  }
  dip3[0] = sx;
  dip3[1] = sy;
  dip3[2] = sz;
}

/* ------------------------------------------------------------------------ */
/* CpuEwaldRecip.cpp:87-147 calcForceEnergy                                  */
/* ------------------------------------------------------------------------ */
double orc_ewald_recip(double sigma, int kcut, double bx, double by, double bz,
                       double ccelec, int n, const double *xyzq, double *fx,
                       double *fy, double *fz) {
  const double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
  const double half_sigmasq = 0.5 * sigma * sigma;
  const double pref = 2.0 * PI_D * ibx * iby * ibz * ccelec;
  for (int i = 0; i < n; i++) fx[i] = fy[i] = fz[i] = 0.0;
  const int kcut2 = kcut * kcut;
  double E = 0.0;
  for (int kz = -kcut; kz <= kcut; kz++) {
    double kzv = 2.0 * PI_D * ((double)kz) * ibz;
    for (int ky = -kcut; ky <= kcut; ky++) {
      double kyv = 2.0 * PI_D * ((double)ky) * iby;
      for (int kx = -kcut; kx <= kcut; kx++) {
        int k2 = kx * kx + ky * ky + kz * kz;
        if ((kx == 0 && ky == 0 && kz == 0) || k2 > kcut2) continue;
        double kxv = 2.0 * PI_D * ((double)kx) * ibx;
        double ksq = kxv * kxv + kyv * kyv + kzv * kzv;
        double re = 0.0, im = 0.0;
        for (int i = 0; i < n; i++) {
          double kr = kxv * xyzq[4 * i] + kyv * xyzq[4 * i + 1] + kzv * xyzq[4 * i + 2];
          re += xyzq[4 * i + 3] * cos(kr);
          im += xyzq[4 * i + 3] * sin(kr);
        }
        double pek = pref * exp(-ksq * half_sigmasq) / ksq;
        for (int i = 0; i < n; i++) {
          double kr = kxv * xyzq[4 * i] + kyv * xyzq[4 * i + 1] + kzv * xyzq[4 * i + 2];
          double s = sin(kr), c = cos(kr);
          double fac = -pek * xyzq[4 * i + 3] * 2.0;
          fx[i] += fac * (re * (-kxv * s) + im * kxv * c);
          fy[i] += fac * (re * (-kyv * s) + im * kyv * c);
          fz[i] += fac * (re * (-kzv * s) + im * kzv * c);
        }
        E += pek * (re * re + im * im);
      }
    }
  }
  return E;
}

/* ------------------------------------------------------------------------ */
/* CpuGSE.cpp:33-44 setParam; rcut as CpuGSEk.cpp:31 and CpuGSEr.cpp:40      */
/* ------------------------------------------------------------------------ */
void orc_gse_derive(const orc_gse_params *p, double *sigma1, double *sigma2,
                    double *rcut1, double *rcut2) {
  double s1 = p->sigma * sqrt(p->kappa);
  double s2 = p->sigma * sqrt(1.0 - 2.0 * p->kappa);
  if (sigma1) *sigma1 = s1;
  if (sigma2) *sigma2 = s2;
  if (rcut1) *rcut1 = p->lambda_sigma1 * sqrt(2.0) * s1;
  if (rcut2) *rcut2 = p->lambda_sigma1 * sqrt(2.0) * s2;
}

void orc_gse_step(const orc_gse_params *p, int mode, int n, const double *xyzq,
                  double *energy, double *energy_k, double *fx, double *fy,
                  double *fz, double *fcx, double *fcy, double *fcz,
                  double *pot, double *vir9, double *rho_out, double *phi_out) {
  const int X = p->X, Y = p->Y, Z = p->Z;
  const size_t sz = (size_t)X * Y * Z;
  double s1, s2, rc1, rc2;
  orc_gse_derive(p, &s1, &s2, &rc1, &rc2);
  double *rho = (double *)malloc(sizeof(double) * sz);
  double *phi = (double *)malloc(sizeof(double) * sz);
  double *rk = (double *)malloc(sizeof(double) * 2 * (size_t)(X / 2 + 1) * Y * Z);
  /* CpuGSEk.cpp:30-32 spreadCharge1 */
  orc_spread(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, rho, 0);
  double Ek = 0.0, virk[9];
  if (mode == 0) {
    /* CpuGSEk.cpp:48-50 -> CpuFFTSolver.cpp:75-95 with G_GSE(sigma2) */
    orc_fft_r2c(X, Y, Z, rho, rk);
    Ek = orc_energy_kspace(0, X, Y, Z, p->ccelec, p->bx, p->by, p->bz, s2, rk,
                           virk);
    orc_greens_mul(0, X, Y, Z, p->bx, p->by, p->bz, s2, rk);
    orc_fft_c2r(X, Y, Z, rk, phi);
  } else {
    /* CpuGSEr.cpp:39-42 on-grid sigma2 spread, then the FFT Poisson solve
     * with the bare 4 pi / k^2 kernel (G_GSE with sigma = 0) */
    double *rhoS = (double *)malloc(sizeof(double) * sz);
    orc_spread_on_grid(X, Y, Z, s2, rc2, p->bx, p->by, p->bz, rho, rhoS);
    orc_fft_solve(0, X, Y, Z, 0.0, p->bx, p->by, p->bz, rhoS, phi);
    free(rhoS);
  }
  /* CpuGSEk.cpp:56-58 / CpuGSEr.cpp:60-62 -> CpuGSE.cpp:86-104 */
  double E = orc_energy(X, Y, Z, p->ccelec, p->bx, p->by, p->bz, rho, phi);
  if (energy) *energy = E;
  if (energy_k) *energy_k = (mode == 0) ? Ek : E;
  /* CpuGSE.cpp:70-80 interpolateForce, x ccelec */
  if (fx || fy || fz) {
    double *ax = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *ay = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *az = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    orc_interpolate(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, phi, ax, ay, az);
    for (int i = 0; i < n; i++) {
      if (fx) fx[i] = ax[i] * p->ccelec;
      if (fy) fy[i] = ay[i] * p->ccelec;
      if (fz) fz[i] = az[i] * p->ccelec;
    }
    free(ax);
    free(ay);
    free(az);
  }
  if (fcx || fcy || fcz || pot || vir9) {
    double *cx = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *cy = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *cz = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *pp = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double vr[9];
    orc_gather_cube(X, Y, Z, s1, rc1, n, xyzq, p->bx, p->by, p->bz, phi, pp,
                    cx, cy, cz, vr);
    for (int i = 0; i < n; i++) {
      if (fcx) fcx[i] = cx[i] * p->ccelec;
      if (fcy) fcy[i] = cy[i] * p->ccelec;
      if (fcz) fcz[i] = cz[i] * p->ccelec;
      if (pot) pot[i] = pp[i] * p->ccelec;
    }
    if (vir9) {
      if (mode == 0) {
        for (int k = 0; k < 9; k++) vir9[k] = virk[k] + vr[k] * p->ccelec;
      } else {
        for (int k = 0; k < 9; k++) vir9[k] = NAN; /* not defined in mode R */
      }
    }
    free(cx);
    free(cy);
    free(cz);
    free(pp);
  }
  if (rho_out) memcpy(rho_out, rho, sizeof(double) * sz);
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * sz);
  free(rho);
  free(phi);
  free(rk);
}

void orc_corrections(double sigma, double bx, double by, double bz,
                     double ccelec, int n, const double *xyzq, double *e_self,
                     double *e_net) {
  double q2 = 0.0, q = 0.0;
  for (int i = 0; i < n; i++) {
    q2 += xyzq[4 * i + 3] * xyzq[4 * i + 3];
    q += xyzq[4 * i + 3];
  }
  if (e_self) *e_self = -ccelec * q2 / sqrt(2.0 * PI_D * sigma * sigma);
  if (e_net) *e_net = -ccelec * PI_D * q * q * sigma * sigma / (bx * by * bz);
}
