// C entry points over the reference's own CPU classes, compiled from the
// reference sources where they lie (oracle/Makefile target `ref`).  TEST
// INFRASTRUCTURE ONLY: it exists so tests can check the oracle restatement
// (gse_oracle.c) against the reference's compiled code.  The output library
// lands in oracle/_ref/, which git ignores.
//
// Wrapped reference routines:
//   CpuGaussCharge<AT,CT>  CpuGaussCharge.h:30-48 (spread, gather, conv,
//                          e-field, dipole, support size)
//   CpuConvolution<T>      CpuConvolution.h:33-37
//   CpuGreensFuncGSE<T>    CpuGreensFuncGSE.h:16-19
//   CpuGreensFuncG2<T>     CpuGreensFuncG2.h:16-19
//   CpuEwaldRecip          CpuEwaldRecip.h:29-31
//   CpuEwaldDirect         CpuEwaldDirect.h:35-43
//   CpuGSEu<AT,CT>         CpuGSEu.h:19-43 (writes rhoG%d.txt into the
//                          working directory at every solve)
//   CpuGSEr<AT,CT>         CpuGSEr.h, with CpuMultiGridSolver (writes
//                          phiCoarsest.txt, rhoCoarsest.txt, resCoarsest.txt)
// CpuFFTSolver / CpuGSEk need FFTW, which this image does not have; they are
// not built.
#include <cstring>

#include "CpuConvolution.h"
#include "CpuEwaldDirect.h"
#include "CpuEwaldRecip.h"
#include "CpuGaussCharge.h"
#include "CpuGreensFuncG2.h"
#include "CpuGreensFuncGSE.h"
#include "CpuGSEr.h"
#include "CpuGSEu.h"

template <typename T>
static void to_xyzq(int n, const double* in, xyzq_t<T>* out) {
  for (int i = 0; i < n; i++) {
    out[i].x = (T)in[4 * i];
    out[i].y = (T)in[4 * i + 1];
    out[i].z = (T)in[4 * i + 2];
    out[i].q = (T)in[4 * i + 3];
  }
}

template <typename T>
static void grid_in(CpuGrid<T>& g, const T* src) {
  std::memcpy(g.getDataPointer(), src, sizeof(T) * g.getSize());
}

template <typename T>
static void grid_out(const CpuGrid<T>& g, T* dst) {
  std::memcpy(dst, g.getDataPointer(), sizeof(T) * g.getSize());
}

template <typename AT, typename CT>
static void spread_t(int X, int Y, int Z, double sigma, double rcut, int n,
                     const double* xyzq, double bx, double by, double bz,
                     CT* rho_out, int normalize) {
  CpuGaussCharge<AT, CT> g(X, Y, Z);
  CpuGrid<CT> rho(X, Y, Z);
  xyzq_t<CT>* p = new xyzq_t<CT>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  g.spreadChargeToGrid(sigma, rcut, n, p, bx, by, bz, rho, normalize != 0);
  grid_out(rho, rho_out);
  delete[] p;
}

template <typename AT, typename CT>
static void interp_t(int X, int Y, int Z, double sigma, double rcut, int n,
                     const double* xyzq, double bx, double by, double bz,
                     const CT* phi_in, AT* fx, AT* fy, AT* fz) {
  CpuGaussCharge<AT, CT> g(X, Y, Z);
  CpuGrid<CT> phi(X, Y, Z);
  grid_in(phi, phi_in);
  xyzq_t<CT>* p = new xyzq_t<CT>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  g.interpolateForce(sigma, rcut, n, p, bx, by, bz, phi, fx, fy, fz);
  delete[] p;
}

extern "C" {

void ref_spread_dd(int X, int Y, int Z, double sigma, double rcut, int n,
                   const double* xyzq, double bx, double by, double bz,
                   double* rho, int normalize) {
  spread_t<double, double>(X, Y, Z, sigma, rcut, n, xyzq, bx, by, bz, rho, normalize);
}

void ref_spread_ff(int X, int Y, int Z, double sigma, double rcut, int n,
                   const double* xyzq, double bx, double by, double bz,
                   float* rho, int normalize) {
  spread_t<float, float>(X, Y, Z, sigma, rcut, n, xyzq, bx, by, bz, rho, normalize);
}

void ref_interpolate_dd(int X, int Y, int Z, double sigma, double rcut, int n,
                        const double* xyzq, double bx, double by, double bz,
                        const double* phi, double* fx, double* fy, double* fz) {
  interp_t<double, double>(X, Y, Z, sigma, rcut, n, xyzq, bx, by, bz, phi, fx, fy, fz);
}

void ref_interpolate_df(int X, int Y, int Z, double sigma, double rcut, int n,
                        const double* xyzq, double bx, double by, double bz,
                        const float* phi, double* fx, double* fy, double* fz) {
  interp_t<double, float>(X, Y, Z, sigma, rcut, n, xyzq, bx, by, bz, phi, fx, fy, fz);
}

void ref_spread_on_grid_dd(int X, int Y, int Z, double sigma, double rcut,
                           double bx, double by, double bz, const double* rho_in,
                           double* rhoS_out) {
  CpuGaussCharge<double, double> g(X, Y, Z);
  CpuGrid<double> rho(X, Y, Z), rhoS(X, Y, Z);
  grid_in(rho, rho_in);
  g.spreadChargeOnGrid(sigma, rcut, bx, by, bz, rho, rhoS);
  grid_out(rhoS, rhoS_out);
}

void ref_efield_dd(int X, int Y, int Z, double bx, double by, double bz,
                   const double* phi_in, double* ex, double* ey, double* ez) {
  CpuGaussCharge<double, double> g(X, Y, Z);
  CpuGrid<double> phi(X, Y, Z), Ex(X, Y, Z), Ey(X, Y, Z), Ez(X, Y, Z);
  grid_in(phi, phi_in);
  g.calcElectricFieldOnGrid(bx, by, bz, phi, Ex, Ey, Ez);
  grid_out(Ex, ex);
  grid_out(Ey, ey);
  grid_out(Ez, ez);
}

void ref_dipole_dd(int X, int Y, int Z, double bx, double by, double bz,
                   const double* rho_in, double* dip3) {
  CpuGaussCharge<double, double> g(X, Y, Z);
  CpuGrid<double> rho(X, Y, Z);
  grid_in(rho, rho_in);
  g.calcDipoleSum(bx, by, bz, rho, dip3[0], dip3[1], dip3[2]);
}

int ref_support_size(int X, int Y, int Z, double rcut, double bx, double by,
                     double bz) {
  CpuGaussCharge<double, double> g(X, Y, Z);
  return g.calcSupportSize(rcut, bx, by, bz);
}

void ref_conv_run_d(int X, int Y, int Z, int Rx, int Ry, int Rz,
                    const double* kx, const double* ky, const double* kz,
                    const double* in, double* out) {
  CpuConvolution<double> c(X, Y, Z);
  CpuGrid<double> a(X, Y, Z), b(X, Y, Z);
  c.setupKernel(Rx, Ry, Rz, kx, ky, kz);
  grid_in(a, in);
  c.run(a, b);
  grid_out(b, out);
}

// kind 0 = CpuGreensFuncGSE, 1 = CpuGreensFuncG2; rhoK holds
// (X/2+1)*2 x Y x Z doubles, multiplied in place
void ref_greens_mul_d(int kind, int X, int Y, int Z, double bx, double by,
                      double bz, double sigma, double* rhoK) {
  CpuGrid<double> g((X / 2 + 1) * 2, Y, Z);
  grid_in(g, rhoK);
  if (kind == 0) {
    CpuGreensFuncGSE<double> f;
    f.mul(X, Y, Z, bx, by, bz, sigma, g);
  } else {
    CpuGreensFuncG2<double> f;
    f.mul(X, Y, Z, bx, by, bz, sigma, g);
  }
  grid_out(g, rhoK);
}

double ref_ewald_recip(double sigma, int kcut, double bx, double by, double bz,
                       double ccelec, int n, const double* xyzq, double* fx,
                       double* fy, double* fz) {
  CpuEwaldRecip e(sigma, kcut, bx, by, bz, ccelec);
  xyzq_t<double>* p = new xyzq_t<double>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  double E = e.calcForceEnergy(n, p, fx, fy, fz);
  delete[] p;
  return E;
}

}  // extern "C"

// CpuEwaldDirect::setupExclusions + calcForceEnergy (CpuEwaldDirect.cpp:31-131)
extern "C" double ref_ewald_direct(double sigma, double rcut, double bx, double by, double bz,
                                   double ccelec, int n, const double* xyzq, int n_iblo14,
                                   const int* iblo14, const int* inb14, double* fx, double* fy,
                                   double* fz) {
  CpuEwaldDirect d(sigma, rcut, bx, by, bz, ccelec);
  if (n_iblo14 > 0) d.setupExclusions(n_iblo14, iblo14, inb14);
  xyzq_t<double>* p = new xyzq_t<double>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  const double e = d.calcForceEnergy(n, p, fx, fy, fz);
  delete[] p;
  return e;
}

// CpuGSEu driven as cpu_gse.cpp:752-760 does: spreadCharge1, solvePoisson,
// calculateEnergy, interpolateForce; phiM copied out
extern "C" double ref_gseu_step(int N, double b, int X, int Y, int Z, double sigma, double kappa,
                                double ls, double ls1, double bx, double by, double bz, int n,
                                const double* xyzq, double* fx, double* fy, double* fz,
                                double* phi) {
  CpuGSEu<double, double> g(N, b, X, Y, Z, sigma, kappa, ls, ls1, bx, by, bz);
  xyzq_t<double>* p = new xyzq_t<double>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  g.spreadCharge1(n, p);
  g.solvePoisson();
  const double e = g.calculateEnergy();
  g.interpolateForce(n, p, fx, fy, fz);
  grid_out(g.getPhi(), phi);
  delete[] p;
  return e;
}

// CpuGSEr driven as cpu_gse.cpp:696-707 does: spreadCharge1, spreadCharge2,
// solvePoisson (200 red-black sweeps), calculateEnergy, interpolateForce
extern "C" double ref_gser_step(int X, int Y, int Z, double sigma, double kappa, double ls,
                                double ls1, double bx, double by, double bz, int n,
                                const double* xyzq, double* fx, double* fy, double* fz,
                                double* rhoS, double* phi) {
  CpuGSEr<double, double> g(X, Y, Z, sigma, kappa, ls, ls1, bx, by, bz);
  xyzq_t<double>* p = new xyzq_t<double>[n > 0 ? n : 1];
  to_xyzq(n, xyzq, p);
  g.spreadCharge1(n, p);
  g.spreadCharge2();
  g.solvePoisson();
  const double e = g.calculateEnergy();
  g.interpolateForce(n, p, fx, fy, fz);
  grid_out(g.getRhoS(), rhoS);
  grid_out(g.getPhi(), phi);
  delete[] p;
  return e;
}
