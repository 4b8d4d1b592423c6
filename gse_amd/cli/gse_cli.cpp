// gse_cli -- the reference driver's front-end (cpu_gse.cpp:38-126) over this
// library: the same options and checks, the same printed lines and output
// files, with every method run on the GPU through include/gse/gse.hpp.
//
//   gse_cli -r r -L L -M M -sigma sigma -order order -seed seed
//           [-test pair|random|file] [-n numCoord] [-xyzq file]
//           [-iblo14 file -inb14 file] [-rcut rcut] [-kcut kcut]
//
//   -test pair (default)  cpu_gse.cpp:638-893 testPair: charges -1 at the
//                         origin and +1 at r (1,1,1)/sqrt(3); GSE-k (energy,
//                         forces, E-field and potential files), SPME order 4,
//                         the exact Ewald reciprocal sum with kcut 32 M / L
//   -test random          cpu_gse.cpp:423-633 testRandom: numCoord charges
//                         (default 100) from srand(seed)/rand() with charges
//                         +-1; GSE-k, SPME(order), Ewald reciprocal (kcut 12)
//                         and direct (rcut 9), force files and rms lines
//   -test file            cpu_gse.cpp:216-418 test48K on a user's xyzq file
//                         (default 48191 charges, L 80, M 80): GSE-k, SPME,
//                         Ewald reciprocal (kcut 64) and direct (rcut 9, with
//                         the iblo14/inb14 exclusion files when given), the
//                         rms lines before and after adding the direct part
//
// M defaults to (int)L as in the reference; the usage line and exit code 1
// follow cpu_gse.cpp:100-103.  The reference's LES runs are not part of
// this library.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "gse/gse.hpp"
#include "gse/io.hpp"

using namespace gse_amd;

namespace {

const double kKappa = 0.4, kLambdaSigma = 3.0, kLambdaSigma1 = 3.0;

void print_forces(const char* tag, const double* fx, const double* fy, const double* fz) {
  std::printf("%s: %.12lf %.12lf %.12lf | %.12lf %.12lf %.12lf\n", tag, fx[0], fy[0], fz[0],
              fx[1], fy[1], fz[1]);
}

// SPME on device particles, forces -dE/dr on the host (gpu_gse.cu:35-163)
double spme(int n, const xyzq_t<double>* h_xyzq, double L, int ngrid, double sigma, int order,
            double* fx, double* fy, double* fz) {
  void* d = nullptr;
  check(gse_device_malloc(&d, sizeof(xyzq_t<double>) * (size_t)(n > 0 ? n : 1)));
  if (n > 0) check(gse_memcpy_to_device(d, h_xyzq, sizeof(xyzq_t<double>) * n));
  double* f = nullptr;
  check(gse_device_malloc((void**)&f, sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)));
  const double recip[9] = {1.0 / L, 0, 0, 0, 1.0 / L, 0, 0, 0, 1.0 / L};
  double e = 0.0;
  {
    CudaPMERecip<double, double> grid(ngrid, ngrid, ngrid, order);
    grid.spread_charge((const xyzq_t<double>*)d, n, recip);
    grid.r2c_fft();
    grid.scalar_sum(recip, 1.0 / (sigma * std::sqrt(2.0)), true, true);
    grid.c2r_fft();
    grid.gather_force((const xyzq_t<double>*)d, n, recip, n, f);
    e = grid.getEnergy();
  }
  std::vector<double> g(3 * (size_t)n);
  if (n > 0) check(gse_memcpy_to_host(g.data(), f, sizeof(double) * 3 * n));
  for (int i = 0; i < n; i++) {  // gpu_gse.cu:145-149
    fx[i] = -g[i];
    fy[i] = -g[n + i];
    fz[i] = -g[2 * n + i];
  }
  check(gse_device_free(d));
  check(gse_device_free(f));
  return e;
}

void testPair(double r, double L, int ngrid, double sigma) {
  const int order = 4;
  xyzq_t<double> xyzq[2];
  const double a = r / std::sqrt(3.0);
  xyzq[0] = {0.0, 0.0, 0.0, -1.0};
  xyzq[1] = {a + 0.0, a + 0.0, a + 0.0, 1.0};
  double fx[2], fy[2], fz[2];

  CpuGSEk<double, double> gsek(ngrid, ngrid, ngrid, sigma, kKappa, kLambdaSigma, kLambdaSigma1, L,
                               L, L);
  gsek.printInfo();
  gsek.spreadCharge1(2, xyzq);
  gsek.solvePoisson();
  const double energy_GSEk = gsek.calculateEnergy();
  std::cout << "energy_GSEk " << energy_GSEk << std::endl;
  CpuGrid<double> Ex(ngrid, ngrid, ngrid), Ey(ngrid, ngrid, ngrid), Ez(ngrid, ngrid, ngrid);
  CpuGaussCharge<double, double> gaussCharge(ngrid, ngrid, ngrid);
  gaussCharge.calcElectricFieldOnGrid(L, L, L, gsek.getPhi(), Ex, Ey, Ez);
  Ex.save("ExGSEk.txt");
  Ey.save("EyGSEk.txt");
  Ez.save("EzGSEk.txt");
  gsek.getPhi().save("phi.txt");
  gsek.interpolateForce(2, xyzq, fx, fy, fz);
  print_forces("GSEk", fx, fy, fz);

  const double energy_SPME = spme(2, xyzq, L, ngrid, sigma, order, fx, fy, fz);
  std::cout << "energy_SPME " << energy_SPME << std::endl;
  print_forces("SPME", fx, fy, fz);

  double refX[2], refY[2], refZ[2];
  CpuEwaldRecip ewald(sigma, (int)(32 * ngrid / L), L, L, L);
  const double energy_ew = ewald.calcForceEnergy(2, xyzq, refX, refY, refZ);
  std::cout << "energy_ew " << energy_ew << std::endl;
  print_forces("EWALD", refX, refY, refZ);
}

void report(const char* tag, int n, const std::vector<double>& rx, const std::vector<double>& ry,
            const std::vector<double>& rz, const std::vector<double>& x,
            const std::vector<double>& y, const std::vector<double>& z, double eref, double e) {
  std::printf("rms(%s) = %e %e\n", tag, Erms(n, rx.data(), ry.data(), rz.data(), x.data(),
                                             y.data(), z.data()),
              std::fabs((eref - e) / eref));
}

void testRandom(int numCoord, double L, int ngrid, double sigma, int order, unsigned seed) {
  std::srand(seed);  // cpu_gse.cpp:431-438
  std::vector<xyzq_t<double>> xyzq(numCoord);
  for (int i = 0; i < numCoord; i++) {
    xyzq[i].x = L * ((double)std::rand() / ((double)RAND_MAX + 1.0));
    xyzq[i].y = L * ((double)std::rand() / ((double)RAND_MAX + 1.0));
    xyzq[i].z = L * ((double)std::rand() / ((double)RAND_MAX + 1.0));
    xyzq[i].q = (double)((i % 2) * 2 - 1);
  }
  const int n = numCoord;
  std::vector<double> gx(n), gy(n), gz(n), sx(n), sy(n), sz(n), ex(n), ey(n), ez(n), dx(n),
      dy(n), dz(n);
  CpuGSEk<double, double> gsek(ngrid, ngrid, ngrid, sigma, kKappa, kLambdaSigma, kLambdaSigma1, L,
                               L, L);
  gsek.printInfo();
  gsek.spreadCharge1(n, xyzq.data());
  gsek.solvePoisson();
  const double energy_GSEk = gsek.calculateEnergy();
  std::cout << "energy_GSEk = " << energy_GSEk << std::endl;
  gsek.interpolateForce(n, xyzq.data(), gx.data(), gy.data(), gz.data());
  saveForce<double>("forceGSEk.txt", n, gx.data(), gy.data(), gz.data());

  const double energy_SPME = spme(n, xyzq.data(), L, ngrid, sigma, order, sx.data(), sy.data(),
                                  sz.data());
  saveForce<double>("forceSPME.txt", n, sx.data(), sy.data(), sz.data());

  CpuEwaldRecip ewr(sigma, 12, L, L, L);
  const double energy_ew = ewr.calcForceEnergy(n, xyzq.data(), ex.data(), ey.data(), ez.data());
  CpuEwaldDirect ewd(sigma, 9.0, L, L, L);
  const double energy_ew_dir = ewd.calcForceEnergy(n, xyzq.data(), dx.data(), dy.data(), dz.data());
  std::cout << "energy_ew = " << energy_ew << std::endl;
  std::cout << "energy_ew_dir = " << energy_ew_dir << std::endl;
  saveForce<double>("forceEW.txt", n, ex.data(), ey.data(), ez.data());
  report("GSEk", n, ex, ey, ez, gx, gy, gz, energy_ew, energy_GSEk);
  report("SPME", n, ex, ey, ez, sx, sy, sz, energy_ew, energy_SPME);
}

void testFile(const char* file, int numCoord, double L, int ngrid, double sigma, int order,
              const char* iblo, const char* inb, double rcut, int kcut) {
  std::vector<xyzq_t<double>> xyzq(numCoord);
  loadArray<double>(4, numCoord, file, (double*)xyzq.data());
  const int n = numCoord;
  std::vector<double> gx(n), gy(n), gz(n), sx(n), sy(n), sz(n), ex(n), ey(n), ez(n), dx(n),
      dy(n), dz(n);
  CpuGSEk<double, double> gsek(ngrid, ngrid, ngrid, sigma, kKappa, kLambdaSigma, kLambdaSigma1, L,
                               L, L);
  gsek.printInfo();
  gsek.spreadCharge1(n, xyzq.data());
  gsek.solvePoisson();
  double energy_GSEk = gsek.calculateEnergy();
  std::cout << "energy_GSEk = " << energy_GSEk << std::endl;
  gsek.interpolateForce(n, xyzq.data(), gx.data(), gy.data(), gz.data());
  saveForce<double>("forceGSEk.txt", n, gx.data(), gy.data(), gz.data());

  double energy_SPME = spme(n, xyzq.data(), L, ngrid, sigma, order, sx.data(), sy.data(),
                            sz.data());
  saveForce<double>("forceSPME.txt", n, sx.data(), sy.data(), sz.data());

  CpuEwaldRecip ewr(sigma, kcut, L, L, L);
  double energy_ew = ewr.calcForceEnergy(n, xyzq.data(), ex.data(), ey.data(), ez.data());
  std::cout << "energy_ew = " << energy_ew << std::endl;
  saveForce<double>("forceEW.txt", n, ex.data(), ey.data(), ez.data());
  CpuEwaldDirect ewd(sigma, rcut, L, L, L);
  if (iblo && inb) {  // cpu_gse.cpp:335-341
    std::vector<int> ib(n);
    loadArray<int>(1, n, iblo, ib.data());
    const int ninb = ib[n - 1];
    std::vector<int> nb(ninb > 0 ? ninb : 1);
    loadArray<int>(1, ninb, inb, nb.data());
    ewd.setupExclusions(n, ib.data(), nb.data());
  }
  const double energy_ew_dir = ewd.calcForceEnergy(n, xyzq.data(), dx.data(), dy.data(), dz.data());
  saveForce<double>("forceEWdir.txt", n, dx.data(), dy.data(), dz.data());
  report("GSEk", n, ex, ey, ez, gx, gy, gz, energy_ew, energy_GSEk);
  report("SPME", n, ex, ey, ez, sx, sy, sz, energy_ew, energy_SPME);
  // cpu_gse.cpp:358-381 adds the direct part to every method
  for (int i = 0; i < n; i++) {
    sx[i] += dx[i];
    sy[i] += dy[i];
    sz[i] += dz[i];
    gx[i] += dx[i];
    gy[i] += dy[i];
    gz[i] += dz[i];
    ex[i] += dx[i];
    ey[i] += dy[i];
    ez[i] += dz[i];
  }
  energy_SPME += energy_ew_dir;
  energy_GSEk += energy_ew_dir;
  energy_ew += energy_ew_dir;
  report("GSEk", n, ex, ey, ez, gx, gy, gz, energy_ew, energy_GSEk);
  report("SPME", n, ex, ey, ez, sx, sy, sz, energy_ew, energy_SPME);
}

int usage() {
  std::cout << "Usage: cpu_gse -r r -L L -M M -sigma sigma -order order -seed seed" << std::endl;
  return 1;
}

}  // namespace

int main(int argc, char* argv[]) {
  double r = 4.0, L = 16.0, sigma = 1.5, rcut = 9.0;
  int M = -1, order = 4, numCoord = -1, kcut = -1;
  unsigned seed = 1234567890;
  std::string test = "pair";
  const char *xyzqFile = nullptr, *iblo = nullptr, *inb = nullptr;
  bool arg_ok = true;
  for (int i = 1; i < argc && arg_ok; i++) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* {
      if (i + 1 >= argc) {
        arg_ok = false;
        return nullptr;
      }
      return argv[++i];
    };
    const char* v = nullptr;
    if (a == "-r") { if ((v = next())) std::sscanf(v, "%lf", &r); }
    else if (a == "-L") { if ((v = next())) std::sscanf(v, "%lf", &L); }
    else if (a == "-M") { if ((v = next())) std::sscanf(v, "%d", &M); }
    else if (a == "-sigma") { if ((v = next())) std::sscanf(v, "%lf", &sigma); }
    else if (a == "-order") { if ((v = next())) std::sscanf(v, "%d", &order); }
    else if (a == "-seed") { if ((v = next())) std::sscanf(v, "%u", &seed); }
    else if (a == "-test") { if ((v = next())) test = v; }
    else if (a == "-n") { if ((v = next())) std::sscanf(v, "%d", &numCoord); }
    else if (a == "-xyzq") { xyzqFile = next(); }
    else if (a == "-iblo14") { iblo = next(); }
    else if (a == "-inb14") { inb = next(); }
    else if (a == "-rcut") { if ((v = next())) std::sscanf(v, "%lf", &rcut); }
    else if (a == "-kcut") { if ((v = next())) std::sscanf(v, "%d", &kcut); }
    else {
      std::cout << "Invalid input parameter " << a << std::endl;
      arg_ok = false;
    }
  }
  if (M < 0) M = (int)L;
  if (!arg_ok || r < 0.0 || L <= 0.0 || M <= 0 || (order != 4 && order != 6 && order != 8) ||
      (test != "pair" && test != "random" && test != "file") || (test == "file" && !xyzqFile))
    return usage();
  try {
    if (test == "pair") {
      testPair(r, L, M, sigma);
    } else if (test == "random") {
      testRandom(numCoord > 0 ? numCoord : 100, L, M, sigma, order, seed);
    } else {
      testFile(xyzqFile, numCoord > 0 ? numCoord : 48191, L, M, sigma, order, iblo, inb, rcut,
               kcut >= 0 ? kcut : 64);
    }
  } catch (const Error& e) {
    std::cerr << e.what() << std::endl;
    return 2;
  }
  return 0;
}
