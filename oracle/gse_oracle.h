/*
 * FP64 CPU oracle for the GSE reciprocal-space path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, loop for loop, the
 * reference algorithm of ap-hynninen/GSE for the k-space GSE pipeline
 * (spread -> [on-grid convolution] -> FFT * G(k) -> inverse FFT -> energy ->
 * gather).  It is used by tests/ as the checker, by __graft_entry__.smoke()
 * as the checker and by bench.py as the timed CPU baseline.  It is never
 * linked into, loaded by or called from the product library (gse_amd/).
 *
 * Every function cites the reference file:line it restates.  FFTW (which the
 * reference calls) is absent from this image, so the 3D transforms are a
 * self-contained mixed-radix FFT with FFTW's conventions: unnormalised,
 * forward exp(-i k x), backward exp(+i k x), r2c output [z][y][x<=X/2].
 *
 * Grids are x-fastest: data[x + (y + z*Y)*X].
 * Particles are AoS records of 4 doubles {x, y, z, q}.
 */
#ifndef GSE_ORACLE_H
#define GSE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* CpuGaussCharge.cpp:28-56 */
int orc_support_size(int X, int Y, int Z, double rcut,
                     double bx, double by, double bz);

/* CpuGaussCharge.cpp:235-384 (spreadChargeToGrid, CT=AT=double) */
void orc_spread(int X, int Y, int Z, double sigma, double rcut,
                int n, const double *xyzq, double bx, double by, double bz,
                double *rho, int normalize);

/* CpuGaussCharge.cpp:390-519 (spreadChargeOnGrid): 3-pass X->Y->Z */
void orc_spread_on_grid(int X, int Y, int Z, double sigma, double rcut,
                        double bx, double by, double bz,
                        const double *rho, double *rhoS);

/* CpuConvolution.cpp:116-151 (run): X pass only, then copy */
void orc_conv_run_x(int X, int Y, int Z, int Rx, const double *kx,
                    const double *in, double *out);

/* Separable 3-pass periodic convolution with per-axis taps, same loop
 * structure as CpuGaussCharge.cpp:448-514. */
void orc_conv_run_xyz(int X, int Y, int Z, int Rx, int Ry, int Rz,
                      const double *kx, const double *ky, const double *kz,
                      const double *in, double *out);

/* CpuGaussCharge.cpp:105-172 (interpolateLoop): sphere support r2<rcut2,
 * result not yet multiplied by ccelec. */
void orc_interpolate(int X, int Y, int Z, double sigma, double rcut,
                     int n, const double *xyzq, double bx, double by, double bz,
                     const double *phi, double *fx, double *fy, double *fz);

/* Build-defined gather over the full (2n+1)^3 cube (the support the spread
 * uses): per-atom potential phi_i (without ccelec), cube force (without
 * ccelec, the exact -dE/dr_i up to ccelec) and the per-system real-space
 * virial moment sum_i sum_p s_ip d_a d_b (9 entries, row-major).  Any output
 * pointer may be NULL. */
void orc_gather_cube(int X, int Y, int Z, double sigma, double rcut,
                     int n, const double *xyzq, double bx, double by, double bz,
                     const double *phi, double *pot, double *fx, double *fy,
                     double *fz, double *vir9);

/* 3D r2c / c2r, FFTW conventions (CpuFFTSolver.cpp:41-44, 61-64). The
 * complex array holds (X/2+1)*Y*Z interleaved (re,im) pairs. c2r does not
 * modify its input. */
void orc_fft_r2c(int X, int Y, int Z, const double *in, double *outc);
void orc_fft_c2r(int X, int Y, int Z, const double *inc, double *out);
/* 1D complex transform of length n (sign -1 forward, +1 backward), in place */
void orc_fft_1d(int n, double *data, int sign);

/* CpuGreensFuncGSE.cpp:9-56 and CpuGreensFuncG2.cpp:9-60; kind 0 = GSE
 * (sigma = 0 gives the plain 4*pi/k^2 Coulomb kernel), kind 1 = G2. */
void orc_greens_mul(int kind, int X, int Y, int Z, double bx, double by,
                    double bz, double sigma, double *rhoK);

/* CpuFFTSolver.cpp:75-95: phi = c2r(G * r2c(rho)), unnormalised */
void orc_fft_solve(int kind, int X, int Y, int Z, double sigma,
                   double bx, double by, double bz,
                   const double *rho, double *phi);

/* CpuGSE.cpp:86-104: E = 0.5*ccelec*hx*hy*hz*sum(rho*phi) */
double orc_energy(int X, int Y, int Z, double ccelec, double bx, double by,
                  double bz, const double *rho, const double *phi);

/* Build-defined: the same energy by Parseval in k-space, plus the k-space
 * part of the virial (9 entries; may be NULL).  rhoK is r2c(rho),
 * unmultiplied.  kind/sigma as orc_greens_mul. */
double orc_energy_kspace(int kind, int X, int Y, int Z, double ccelec,
                         double bx, double by, double bz, double sigma,
                         const double *rhoK, double *vir9);

/* CpuGaussCharge.cpp:62-99 */
void orc_efield(int X, int Y, int Z, double bx, double by, double bz,
                const double *phi, double *ex, double *ey, double *ez);

/* CpuGaussCharge.cpp:189-226 */
void orc_dipole_sum(int X, int Y, int Z, double bx, double by, double bz,
                    const double *rho, double *dip3);

/* CpuEwaldRecip.cpp:87-147 */
double orc_ewald_recip(double sigma, int kcut, double bx, double by, double bz,
                       double ccelec, int n, const double *xyzq,
                       double *fx, double *fy, double *fz);

/* CpuGSE.cpp:33-44 plus the rcut choice of CpuGSEk.cpp:31 / CpuGSE.cpp:72 */
typedef struct {
  int X, Y, Z;
  double sigma, kappa, lambda_sigma, lambda_sigma1;
  double bx, by, bz;
  double ccelec;
} orc_gse_params;

void orc_gse_derive(const orc_gse_params *p, double *sigma1, double *sigma2,
                    double *rcut1, double *rcut2);

/* The full GSE-k step (CpuGSEk.cpp:30-58 driven as in cpu_gse.cpp:719-735)
 * in mode 0 = K (spread sigma1, FFT with G_GSE(sigma2)) or mode 1 = R
 * (spread sigma1, on-grid sigma2 convolution, FFT with G_GSE(0), energy from
 * the sigma1 grid as CpuGSEr.cpp:60-62).  Outputs (any may be NULL):
 *   energy      : real-space energy, CpuGSE.cpp:86-104
 *   energy_k    : Parseval energy from the k-pass (mode K only, else = energy)
 *   fx,fy,fz    : reference sphere-support force times ccelec (CpuGSE.cpp:70-80)
 *   fcx,fcy,fcz : cube-support force times ccelec (= -dE/dr exactly)
 *   pot         : per-atom potential phi_i = dE/dq_i (ccelec included)
 *   vir9        : virial W_ab = -dE/d(eps_ab), atoms and grid co-deforming
 *   rho_out, phi_out : the sigma1 grid and the potential grid            */
void orc_gse_step(const orc_gse_params *p, int mode, int n, const double *xyzq,
                  double *energy, double *energy_k,
                  double *fx, double *fy, double *fz,
                  double *fcx, double *fcy, double *fcz,
                  double *pot, double *vir9,
                  double *rho_out, double *phi_out);

/* Closed-form corrections (build-defined, CpuEwaldDirect.cpp:124-128 for the
 * self term): self energy -ccelec*sum(q^2)/sqrt(2*pi*sigma^2) and neutralising
 * background term -ccelec*pi*Q^2*sigma^2/V. */
void orc_corrections(double sigma, double bx, double by, double bz,
                     double ccelec, int n, const double *xyzq,
                     double *e_self, double *e_net);

/* ---- GSE-u and GSE-r (CpuGSEu.cpp, CpuGSEr.cpp, CpuMultiGridSolver.cpp),
 * gse_ru_oracle.c ---- */
/* nsweep red-black Gauss-Seidel sweeps of the HERMITIAN stencil on phi (in
 * place), CpuMultiGridSolver.cpp:204-290; rho is the source (-4 pi rhoS) */
void orc_rbgs_hermitian(int X, int Y, int Z, double bx, double by, double bz,
                        const double *rho, double *phi, int nsweep);
/* -4 pi B^/(A^ XYZ) of half-spectrum index (ix, iy, iz), 0 at k = 0 */
double orc_mehrstellen_g(int X, int Y, int Z, double bx, double by, double bz, int ix, int iy,
                         int iz);
/* zero-mean exact solution of the HERMITIAN system for source -4 pi rhoS */
void orc_mehrstellen_solve(int X, int Y, int Z, double bx, double by, double bz,
                           const double *rhoS, double *phi);
/* GSE-r step; nsweep > 0: the reference's relaxation with that many sweeps
 * from phi = 0, nsweep = 0: the exact (zero-mean) solution */
void orc_gser_step(const orc_gse_params *p, int nsweep, int n, const double *xyzq,
                   double *energy, double *fx, double *fy, double *fz, double *rho_out,
                   double *rhoS_out, double *phi_out);
/* GSE-u step with N terms and base b */
void orc_gseu_step(int N, double b, const orc_gse_params *p, int n, const double *xyzq,
                   double *energy, double *fx, double *fy, double *fz, double *phi_out);

/* ---- Ewald direct space + self energy (CpuEwaldDirect.cpp:31-131), ewald_oracle.c ----
 * Returns E_direct + E_self; fx, fy, fz receive dE/dr_i (the reference's
 * sign); exclusions in the reference's iblo14/inb14 form (n_iblo14 = 0: none). */
double orc_ewald_direct(double sigma, double rcut, double boxx, double boxy, double boxz,
                        double ccelec, int n, const double *xyzq, int n_iblo14,
                        const int *iblo14, const int *inb14, double *fx, double *fy,
                        double *fz);

/* ---- SPME reciprocal comparator (cuda/CudaPMERecip.cu), spme_oracle.c ---- */
/* B-spline weights theta[0..order-1] and their derivatives at fraction w
 * (calc_theta_dtheta, CudaPMERecip.cu:266-327) */
void orc_pme_theta(int order, double w, double *theta, double *dtheta);
/* 1/|b(k)|^2 B-spline moduli of one axis (calc_prefac/dftmod, :3385-3478) */
void orc_pme_prefac(int order, int nfft, double *prefac);
/* Q = sum_i q_i theta theta theta (spread_charge_ortho_*, :436-934) */
void orc_pme_spread(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                    double by, double bz, double *Q);
/* scalar_sum (:946-1256) in place on the r2c half spectrum; energy and the
 * 3x3 virial (row-major) may be NULL */
void orc_pme_scalar_sum(int X, int Y, int Z, int order, double bx, double by, double bz,
                        double kappa, double ccelec, double *Qk, double *energy,
                        double *vir9);
/* energy gradient dE/dr from the convolved grid (gather_force, :1612-1890) */
void orc_pme_gather(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                    double by, double bz, double ccelec, const double *Qc, double *gx,
                    double *gy, double *gz);
/* full step (gpu_gse.cu:61-73): energy, virial, forces -dE/dr, the charge grid */
void orc_pme_step(int X, int Y, int Z, int order, int n, const double *xyzq, double bx,
                  double by, double bz, double kappa, double ccelec, double *energy,
                  double *vir9, double *fx, double *fy, double *fz, double *Q_out);

#ifdef __cplusplus
}
#endif
#endif
