// C++ caller of the drop-in classes (include/gse/gse.hpp) written the way
// cpu_gse.cpp drives the reference: the testPair GSE-k sequence
// (cpu_gse.cpp:638-736), the TEST_CHARGE_SPREAD identity through
// CpuFFTSolver<CpuGreensFuncG2> (cpu_gse.cpp:752-794), and every other
// CpuGaussCharge / CpuConvolution / CpuFFTSolver method, and the reference's
// GPU harness sequences gpu_conv.cu:27-76 and gpu_gauss.cu:52-108 through
// CudaGrid / CudaConvolution / CudaGaussCharge.  Each result is
// checked against the FP64 oracle (oracle/liboracle_gse.so, test
// infrastructure).  Prints one line per check and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>

#include "gse/gse.hpp"
#include "../../oracle/gse_oracle.h"

using namespace gse_amd;

static int g_fail = 0;

static double maxrel(const double* a, const double* b, size_t n) {
  double m = 0.0, d = 0.0;
  for (size_t i = 0; i < n; i++) {
    m = std::fmax(m, std::fabs(b[i]));
    d = std::fmax(d, std::fabs(a[i] - b[i]));
  }
  return m > 0 ? d / m : d;
}

template <typename T>
static std::vector<double> as_double(const T* p, size_t n) {
  return std::vector<double>(p, p + n);
}

static void report(const char* what, double err, double tol) {
  const bool ok = err <= tol;
  std::printf("%-44s %.3e (tol %.0e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
  if (!ok) g_fail++;
}

// splitmix64 -> uniform [0,1)
static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double urand() {
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int main() {
  // ---------------------------------------------------------------- testPair
  {
    const double r = 1.0, L = 20.0, sigma = 2.121320344;
    const int M = 20;
    xyzq_t<double> xyzq[2];
    const double a = r / std::sqrt(3.0);  // cpu_gse.cpp:689-698
    xyzq[0] = {0.0, 0.0, 0.0, -1.0};
    xyzq[1] = {a, a, a, 1.0};
    double fx[2], fy[2], fz[2];
    CpuGSEk<double, double> gsek(M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L);
    gsek.printInfo();
    gsek.spreadCharge1(2, xyzq);
    gsek.solvePoisson();
    const double e = gsek.calculateEnergy();
    CpuGrid<double> Ex(M, M, M), Ey(M, M, M), Ez(M, M, M);
    CpuGaussCharge<double, double> gc(M, M, M);
    gc.calcElectricFieldOnGrid(L, L, L, gsek.getPhi(), Ex, Ey, Ez);
    gsek.interpolateForce(2, xyzq, fx, fy, fz);
    std::printf("energy_GSEk %.12f\nGSEk: %.12f %.12f %.12f | %.12f %.12f %.12f\n", e, fx[0],
                fy[0], fz[0], fx[1], fy[1], fz[1]);

    orc_gse_params p = {M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L, 332.0716};
    double oe, oek, ofx[2], ofy[2], ofz[2];
    std::vector<double> orho(M * M * M), ophi(M * M * M);
    orc_gse_step(&p, 0, 2, &xyzq[0].x, &oe, &oek, ofx, ofy, ofz, nullptr, nullptr, nullptr,
                 nullptr, nullptr, orho.data(), ophi.data());
    report("pair CpuGSEk energy", std::fabs(e - oe) / std::fabs(oe), 1e-10);
    double f[6] = {fx[0], fy[0], fz[0], fx[1], fy[1], fz[1]};
    double of[6] = {ofx[0], ofy[0], ofz[0], ofx[1], ofy[1], ofz[1]};
    report("pair CpuGSEk interpolateForce", maxrel(f, of, 6), 1e-10);
    report("pair CpuGSEk getRhoS", maxrel(gsek.getRhoS().getDataPointer(), orho.data(), orho.size()), 1e-10);
    report("pair CpuGSEk getPhi", maxrel(gsek.getPhi().getDataPointer(), ophi.data(), ophi.size()), 1e-10);
    std::vector<double> ex(M * M * M), ey(M * M * M), ez(M * M * M);
    orc_efield(M, M, M, L, L, L, ophi.data(), ex.data(), ey.data(), ez.data());
    report("pair calcElectricFieldOnGrid Ex", maxrel(Ex.getDataPointer(), ex.data(), ex.size()), 1e-10);
    report("pair calcElectricFieldOnGrid Ez", maxrel(Ez.getDataPointer(), ez.data(), ez.size()), 1e-10);

    // TEST_CHARGE_SPREAD: sigma_t spread vs G2 back-and-forth FFT of rhoS
    const double s1 = gsek.getSigma1(), s2 = gsek.getSigma2();
    const double st = std::sqrt(s1 * s1 + s2 * s2);
    CpuGrid<double> rhoS_ref(M, M, M), rhoS_k(M, M, M);
    gc.spreadChargeToGrid(st, 3.0 * std::sqrt(2.0) * st, 2, xyzq, L, L, L, rhoS_ref);
    CpuGreensFunc<double>* g2 = new CpuGreensFuncG2<double>();
    CpuFFTSolver<double> solver(*g2, M, M, M);
    solver.run(rhoS_k, gsek.getRhoS(), s2, L, L, L);
    delete g2;
    CpuGrid<double> res(M, M, M);
    res.copy(rhoS_k);
    res.sub(rhoS_ref);
    const double re = res.maxAbsValue() / rhoS_ref.maxAbsValue();
    std::printf("RE|rhoS_k - rhoS_ref| = %e\n", re);
    report("pair sigma1*sigma2 = sigma_t spreading identity", re, 2e-3);
    std::vector<double> ok_ref(M * M * M), ok_k(M * M * M);
    orc_spread(M, M, M, st, 3.0 * std::sqrt(2.0) * st, 2, &xyzq[0].x, L, L, L, ok_ref.data(), 0);
    orc_fft_solve(1, M, M, M, s2, L, L, L, orho.data(), ok_k.data());
    report("pair spreadChargeToGrid(sigma_t)", maxrel(rhoS_ref.getDataPointer(), ok_ref.data(), ok_ref.size()), 1e-10);
    report("pair CpuFFTSolver<G2>::run", maxrel(rhoS_k.getDataPointer(), ok_k.data(), ok_k.size()), 1e-10);
  }

  // ----------------------------------------- random box, every class method
  {
    const int X = 18, Y = 16, Z = 20, N = 300;
    const double bx = 17.5, by = 16.2, bz = 20.4, sigma = 1.34164, rcut = 3.0 * std::sqrt(2.0) * sigma;
    std::vector<xyzq_t<double>> p(N);
    for (int i = 0; i < N; i++)
      p[i] = {urand() * bx, urand() * by, urand() * bz, (i & 1) ? 0.417 : -0.417};
    const double* xq = &p[0].x;
    CpuGaussCharge<double, double> gc(X, Y, Z);
    CpuGrid<double> rho(X, Y, Z), rhoS(X, Y, Z);
    const size_t np = (size_t)X * Y * Z;
    std::vector<double> o(np), o2(np);

    report("calcSupportSize", std::fabs((double)gc.calcSupportSize(rcut, bx, by, bz) -
                                        orc_support_size(X, Y, Z, rcut, bx, by, bz)), 0.0);
    for (int norm = 0; norm < 2; norm++) {
      gc.spreadChargeToGrid(sigma, rcut, N, p.data(), bx, by, bz, rho, norm != 0);
      orc_spread(X, Y, Z, sigma, rcut, N, xq, bx, by, bz, o.data(), norm);
      report(norm ? "spreadChargeToGrid(normalize=true)" : "spreadChargeToGrid",
             maxrel(rho.getDataPointer(), o.data(), np), 1e-10);
    }
    gc.spreadChargeOnGrid(0.9487, 4.02, bx, by, bz, rho, rhoS);
    orc_spread_on_grid(X, Y, Z, 0.9487, 4.02, bx, by, bz, o.data(), o2.data());
    report("spreadChargeOnGrid", maxrel(rhoS.getDataPointer(), o2.data(), np), 1e-10);

    std::vector<double> fx(N), fy(N), fz(N), ox(N), oy(N), oz(N);
    gc.interpolateForce(sigma, rcut, N, p.data(), bx, by, bz, rhoS, fx.data(), fy.data(), fz.data());
    orc_interpolate(X, Y, Z, sigma, rcut, N, xq, bx, by, bz, o2.data(), ox.data(), oy.data(), oz.data());
    report("interpolateForce x", maxrel(fx.data(), ox.data(), N), 1e-10);
    report("interpolateForce z", maxrel(fz.data(), oz.data(), N), 1e-10);

    double d[3], od[3];
    gc.calcDipoleSum(bx, by, bz, rho, d[0], d[1], d[2]);
    orc_dipole_sum(X, Y, Z, bx, by, bz, o.data(), od);
    report("calcDipoleSum", maxrel(d, od, 3), 1e-10);

    // CpuConvolution: X pass only, as the reference run()
    const int R = 3;
    double kx[2 * R + 1], ky[2 * R + 1], kz[2 * R + 1];
    for (int i = -R; i <= R; i++) {
      kx[i + R] = std::exp(-0.3 * i * i);
      ky[i + R] = std::exp(-0.2 * i * i);
      kz[i + R] = std::exp(-0.1 * i * i);
    }
    CpuConvolution<double> conv(X, Y, Z);
    conv.setupKernel(R, R, R, kx, ky, kz);
    CpuGrid<double> out(X, Y, Z);
    conv.run(rho, out);
    orc_conv_run_x(X, Y, Z, R, kx, o.data(), o2.data());
    report("CpuConvolution::run (X pass)", maxrel(out.getDataPointer(), o2.data(), np), 1e-10);
    conv.runXYZ(rho, out);
    orc_conv_run_xyz(X, Y, Z, R, R, R, kx, ky, kz, o.data(), o2.data());
    report("CpuConvolution::runXYZ", maxrel(out.getDataPointer(), o2.data(), np), 1e-10);

    // CpuFFTSolver with the built-in GSE kernel and with a caller-defined
    // kernel (host mul between the device transforms)
    struct MyGSE : public CpuGreensFunc<double> {
      void mul(const int nX, const int nY, const int nZ, const double bx, const double by,
               const double bz, const double sigma, CpuGrid<double>& rhoK) override {
        orc_greens_mul(0, nX, nY, nZ, bx, by, bz, sigma, rhoK.getDataPointer());
      }
    } mine;
    CpuGreensFuncGSE<double> gse;
    CpuFFTSolver<double> s_dev(gse, X, Y, Z), s_host(mine, X, Y, Z);
    CpuGrid<double> phi1(X, Y, Z), phi2(X, Y, Z);
    s_dev.run(phi1, rho, 0.9487, bx, by, bz);
    s_host.run(phi2, rho, 0.9487, bx, by, bz);
    orc_fft_solve(0, X, Y, Z, 0.9487, bx, by, bz, o.data(), o2.data());
    report("CpuFFTSolver<GSE>::run (device k-pass)", maxrel(phi1.getDataPointer(), o2.data(), np), 1e-10);
    report("CpuFFTSolver<caller kernel>::run (host mul)", maxrel(phi2.getDataPointer(), o2.data(), np), 1e-10);
  }

  // ----------------------------------- FP32 and mixed instantiations (Erms)
  {
    const int M = 32, N = 3000;
    const double L = 31.0, sigma = 2.12;
    std::vector<xyzq_t<double>> p(N);
    for (int i = 0; i < N; i++)
      p[i] = {urand() * L, urand() * L, urand() * L, (i % 3 == 0) ? -0.834 : 0.417};
    orc_gse_params op = {M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L, 332.0716};
    std::vector<double> ofx(N), ofy(N), ofz(N);
    double oe, oek;
    orc_gse_step(&op, 0, N, &p[0].x, &oe, &oek, ofx.data(), ofy.data(), ofz.data(), nullptr,
                 nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    auto erms = [&](const std::vector<double>& fx, const std::vector<double>& fy,
                    const std::vector<double>& fz) {
      double num = 0, den = 0;
      for (int i = 0; i < N; i++) {
        num += (fx[i] - ofx[i]) * (fx[i] - ofx[i]) + (fy[i] - ofy[i]) * (fy[i] - ofy[i]) +
               (fz[i] - ofz[i]) * (fz[i] - ofz[i]);
        den += ofx[i] * ofx[i] + ofy[i] * ofy[i] + ofz[i] * ofz[i];
      }
      return std::sqrt(num / den);
    };
    {
      std::vector<xyzq_t<float>> pf(N);
      for (int i = 0; i < N; i++) pf[i] = {(float)p[i].x, (float)p[i].y, (float)p[i].z, (float)p[i].q};
      // the oracle sees the same (float-rounded) coordinates
      std::vector<xyzq_t<double>> pr(N);
      for (int i = 0; i < N; i++) pr[i] = {pf[i].x, pf[i].y, pf[i].z, pf[i].q};
      orc_gse_step(&op, 0, N, &pr[0].x, &oe, &oek, ofx.data(), ofy.data(), ofz.data(), nullptr,
                   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
      CpuGSEk<float, float> g(M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L);
      std::vector<float> fx(N), fy(N), fz(N);
      g.spreadCharge1(N, pf.data());
      g.solvePoisson();
      const float e = g.calculateEnergy();
      g.interpolateForce(N, pf.data(), fx.data(), fy.data(), fz.data());
      report("CpuGSEk<float,float> 3k Erms", erms(as_double(fx.data(), N), as_double(fy.data(), N),
                                                  as_double(fz.data(), N)), 1e-5);
      report("CpuGSEk<float,float> 3k energy", std::fabs(e - oe) / std::fabs(oe), 1e-5);
      CpuGSEk<double, float> gm(M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L);
      std::vector<double> dx(N), dy(N), dz(N);
      gm.spreadCharge1(N, pf.data());
      gm.solvePoisson();
      gm.interpolateForce(N, pf.data(), dx.data(), dy.data(), dz.data());
      report("CpuGSEk<double,float> 3k Erms", erms(dx, dy, dz), 1e-5);
    }
  }

  // ------------------- gpu_conv.cu:27-76 through CudaGrid / CudaConvolution
  {
    const int n = 128, R = 5;
    const double sigma = 2.12;
    const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
    const double pref = std::pow(2 * pi_dbl * sigma * sigma, -3.0 / 2.0);
    std::vector<double> kx(2 * R + 1), ky(2 * R + 1), kz(2 * R + 1);
    for (int ii = -R; ii <= R; ii++) {
      kx[ii + R] = std::exp(-(ii * ii) * inv_2sigmasq);
      ky[ii + R] = std::exp(-(ii * ii) * inv_2sigmasq);
      kz[ii + R] = pref * std::exp(-(ii * ii) * inv_2sigmasq);
    }
    CpuGrid<double> cpuRho(n, n, n), cpuRhoXYZ(n, n, n);
    CudaGrid<double> cudaRho(n, n, n), cudaRhoXYZ(n, n, n);
    double* d = cpuRho.getDataPointer();
    for (int i = 0; i < n * n * n; i++) d[i] = urand();
    cudaRho.copy(cpuRho);
    CudaConvolution<double> cudaConv(n, n, n);
    cudaConv.setupKernel(R, R, R, kx.data(), ky.data(), kz.data());
    cudaRhoXYZ.clear();
    cudaConv.run(cudaRho, cudaRhoXYZ);
    cpuRhoXYZ.copy(cudaRhoXYZ);  // CpuGrid::copy(const CudaGrid&), device -> host
    const size_t np = (size_t)n * n * n;
    std::vector<double> o(np);
    orc_conv_run_x(n, n, n, R, kx.data(), d, o.data());
    report("CudaConvolution::run n=128 R=5 (X pass)", maxrel(cpuRhoXYZ.getDataPointer(), o.data(), np), 1e-10);
    // CudaGrid::copy(const CudaGrid&) and clear()
    CudaGrid<double> dup(n, n, n);
    dup.copy(cudaRhoXYZ);
    CpuGrid<double> back(n, n, n);
    back.copy(dup);
    report("CudaGrid::copy(CudaGrid) device -> device", maxrel(back.getDataPointer(), o.data(), np), 0.0);
    dup.clear();
    back.copy(dup);
    report("CudaGrid::clear", back.maxAbsValue(), 0.0);
    // the full separable form on device grids, FP32 with compensated sums
    CpuGrid<float> rf(n, n, n), of(n, n, n);
    for (size_t i = 0; i < np; i++) rf.getDataPointer()[i] = (float)d[i];
    CudaGrid<float> cf(n, n, n), cfo(n, n, n);
    cf.copy(rf);
    std::vector<float> kxf(kx.begin(), kx.end()), kyf(ky.begin(), ky.end()), kzf(kz.begin(), kz.end());
    CudaConvolution<float> convf(n, n, n);
    convf.setupKernel(R, R, R, kxf.data(), kyf.data(), kzf.data());
    convf.runXYZ(cf, cfo);
    of.copy(cfo);
    std::vector<double> in32(np), k32x(kxf.begin(), kxf.end()), k32y(kyf.begin(), kyf.end()),
        k32z(kzf.begin(), kzf.end());
    for (size_t i = 0; i < np; i++) in32[i] = rf.getDataPointer()[i];
    orc_conv_run_xyz(n, n, n, R, R, R, k32x.data(), k32y.data(), k32z.data(), in32.data(), o.data());
    report("CudaConvolution<float>::runXYZ n=128", maxrel(as_double(of.getDataPointer(), np).data(), o.data(), np), 1e-6);
  }

  // ----------------- gpu_gauss.cu:52-108 through CudaGaussCharge (device xyzq)
  {
    // the reference reads data/xyzq.txt (23,558 charges, box 62.23); the
    // same count and box with charges of alternating sign stand in for it
    const double sigma = 2.12132034355964 / std::sqrt(2.0), lambdaSigma = 3.0;
    const int n = 64, numCoord = 23558;
    const double box = 62.23;
    std::vector<xyzq_t<double>> h(numCoord);
    for (int i = 0; i < numCoord; i++)
      h[i] = {urand() * box, urand() * box, urand() * box, (i & 1) ? 0.5 : -0.5};
    std::vector<xyzq_t<float>> hf(numCoord);
    for (int i = 0; i < numCoord; i++)
      hf[i] = {(float)h[i].x, (float)h[i].y, (float)h[i].z, (float)h[i].q};
    void* d_xyzq = nullptr;
    check(gse_device_malloc(&d_xyzq, sizeof(xyzq_t<double>) * numCoord));
    check(gse_memcpy_to_device(d_xyzq, h.data(), sizeof(xyzq_t<double>) * numCoord));
    void* d_xyzqf = nullptr;
    check(gse_device_malloc(&d_xyzqf, sizeof(xyzq_t<float>) * numCoord));
    check(gse_memcpy_to_device(d_xyzqf, hf.data(), sizeof(xyzq_t<float>) * numCoord));
    const double rcut = lambdaSigma * std::sqrt(2.0) * sigma;
    const size_t np = (size_t)n * n * n;
    std::vector<double> o(np);
    orc_spread(n, n, n, sigma, rcut, numCoord, &h[0].x, box, box, box, o.data(), 0);
    {
      CudaGaussCharge<double, double> cudaGaussCharge(n, n, n);
      CudaGrid<double> cudaRho(n, n, n);
      cudaGaussCharge.spreadChargeToGrid(sigma, rcut, numCoord, (const xyzq_t<double>*)d_xyzq,
                                         box, box, box, cudaRho);
      CpuGrid<double> cpuRho(n, n, n);
      cpuRho.copy(cudaRho);
      report("CudaGaussCharge<double,double> 23558 charges", maxrel(cpuRho.getDataPointer(), o.data(), np), 1e-10);
    }
    {
      std::vector<xyzq_t<double>> hr(numCoord);
      for (int i = 0; i < numCoord; i++) hr[i] = {hf[i].x, hf[i].y, hf[i].z, hf[i].q};
      orc_spread(n, n, n, sigma, rcut, numCoord, &hr[0].x, box, box, box, o.data(), 0);
      CudaGaussCharge<float, float> cudaGaussCharge(n, n, n);
      CudaGrid<float> cudaRho(n, n, n);
      cudaGaussCharge.spreadChargeToGrid(sigma, rcut, numCoord, (const xyzq_t<float>*)d_xyzqf,
                                         box, box, box, cudaRho);
      CpuGrid<float> cpuRho(n, n, n);
      cpuRho.copy(cudaRho);
      report("CudaGaussCharge<float,float> 23558 charges",
             maxrel(as_double(cpuRho.getDataPointer(), np).data(), o.data(), np), 2e-6);
    }
    check(gse_device_free(d_xyzq));
    check(gse_device_free(d_xyzqf));
  }

  // ------------------------------------------------------------ error path
  try {
    CpuGSEk<double, double> bad(16, 16, 16, 2.0, 0.7, 3.0, 3.0, 16.0, 16.0, 16.0);
    report("kappa > 0.5 throws gse_amd::Error", 1.0, 0.0);
  } catch (const Error& e) {
    std::printf("expected error: %s\n", e.what());
    report("kappa > 0.5 throws gse_amd::Error", 0.0, 0.0);
  }

  std::printf(g_fail ? "FAILED %d checks\n" : "ALL PASSED\n", g_fail);
  return g_fail ? 1 : 0;
}
