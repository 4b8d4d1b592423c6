// Internal host-side declarations shared by the C-ABI layer and the kernel
// translation units.  Nothing here is part of the public boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gse {

// Geometry of one periodic grid and the Gaussian support on it.  The spacing
// and its inverse are derived exactly as CpuGaussCharge.cpp:244-252 does:
// h = box/M, inv_h = 1.0/h, n = (int)ceil(rcut/h).
struct Geom {
  int X, Y, Z;
  double hx, hy, hz;
  double ihx, ihy, ihz;
  int nx, ny, nz;
};

Geom make_geom(int X, int Y, int Z, double bx, double by, double bz, double rcut);

// Cell binning of particles by the wrapped origin of their support cube:
// bin (bx,by,bz) holds the particles whose cube origin o = wrap(round(x/h)-n)
// lies in [b*B, (b+1)*B) on each axis.  The brick a bin touches is
// (B+2n) wide per axis.
struct BinShape {
  int Bx, By, Bz;     // bin edge in grid cells
  int nbx, nby, nbz;  // bins per axis
  int nbins;
  int Lx, Ly, Lz;     // brick edge = B + 2n
  int direct;         // 1: the brick does not fit in LDS; one bin, global atomics
  size_t brick_bytes(size_t elem) const { return (size_t)Lx * Ly * Lz * elem; }
};

BinShape choose_bins(const Geom& g, size_t elem_bytes, size_t lds_budget, int pref_edge);

// Device-side fixed-point parameters of one spread
struct FixedParams {
  double scale;      // 2^e
  double inv_scale;  // 2^-e
  unsigned long long qmax_bits;  // bits of max |q| (as double)
  int invalid;       // number of rejected particles (non-finite / out of range)
  int pad;
};

// Binning scratch owned by a handle
struct Binning {
  int n_cap = 0;
  int bins_cap = 0;
  int* counts = nullptr;     // [bins_cap + 1], exclusive starts after scan
  int* atom_bin = nullptr;   // [n_cap]
  int* atom_slot = nullptr;  // [n_cap]
  int* perm = nullptr;       // [n_cap] sorted slot -> original index
  void* sorted = nullptr;    // [n_cap] 4-vectors of the compute type
  FixedParams* fp = nullptr;
};

// ---- kernel launchers (defined in the .hip units) ------------------------
// ct_double: particle/grid compute type is double (else float)
hipError_t launch_bin(bool ct_double, const Geom& g, const BinShape& b, int n,
                      const void* xyzq, Binning& w, int images, double pref,
                      hipStream_t s);

hipError_t launch_spread(bool ct_double, const Geom& g, const BinShape& b,
                         double sigma, bool normalize, Binning& w,
                         unsigned long long* grid64, hipStream_t s);

// Largest per-contribution weight factor of a spread (1 without normalize;
// with normalize each axis weight is at most sqrt(2 pi) s exp(h^2/8s^2)).
double spread_weight_bound(const Geom& g, double sigma, bool normalize);

hipError_t launch_fixed_to_grid(bool ct_double, const unsigned long long* grid64,
                                void* out, size_t count, const FixedParams* fp,
                                bool clear, hipStream_t s);

hipError_t launch_fill_u64(unsigned long long* p, size_t count, hipStream_t s);
hipError_t launch_fill(bool dbl, void* p, size_t count, hipStream_t s);

// Gather options (bit mask)
enum GatherFlags : int {
  GF_SPHERE = 1,  // reference force: sphere support r2 < rcut2
  GF_CUBE = 2,    // exact gradient: full cube support
  GF_POT = 4,     // per-particle potential
  GF_VIRIAL = 8,  // per-particle virial moments (6, upper triangle)
};

struct GatherOut {
  void* fx = nullptr;  // sphere force (AT)
  void* fy = nullptr;
  void* fz = nullptr;
  void* cx = nullptr;  // cube force (AT)
  void* cy = nullptr;
  void* cz = nullptr;
  void* pot = nullptr;     // potential (AT)
  double* vir6 = nullptr;  // [6][vir_stride] per-particle virial moments
  int vir_stride = 0;
  double fscale = 1.0;     // multiplies the forces (e.g. ccelec)
  double pscale = 1.0;     // multiplies the potential
  double vscale = 1.0;     // multiplies the virial moments
  const double* dscale = nullptr;  // optional device scalar on forces and potential
};

hipError_t launch_gather(bool ct_double, bool at_double, const Geom& g,
                         const BinShape& b, double sigma, double rcut,
                         int flags, const Binning& w, int n, const void* phi,
                         const GatherOut& o, hipStream_t s);

// k-space
hipError_t launch_greens_table(bool dbl, int kind, int X, int Y, int Z,
                               double bx, double by, double bz, double sigma,
                               void* table, hipStream_t s);

// In-place multiply of the half spectrum by the table, FP64 partial sums of
// the energy (and the 6 k-space virial moments when vir is set) per block.
constexpr int kKpassBlocks = 512;
hipError_t launch_kpass(bool dbl, int kind, int X, int Y, int Z, double bx,
                        double by, double bz, double sigma, void* spec,
                        const void* table, bool vir, double* partials,
                        hipStream_t s);
// partials [kKpassBlocks][7] -> out[0] = E, out[1..9] = W (3x3) with
// prefactor 0.5*ccelec*vcell; adds `extra9` (device, may be null) into W.
hipError_t launch_kpass_finalize(const double* partials, double pref, double* out,
                                 bool vir, hipStream_t s);

// real-space reductions
constexpr int kRedBlocks = 512;
hipError_t launch_dot(bool dbl, const void* a, const void* b, size_t n,
                      double* partials, hipStream_t s);
hipError_t launch_reduce_partials(const double* partials, int nblocks, int ncomp, int stride,
                                  double scale, double* out, hipStream_t s);
// adds the 6 symmetric moments summed over [nblocks][6] partials into a 3x3
hipError_t launch_add_sym6(const double* partials, int nblocks, double* vir9, hipStream_t s);
// out[0] = -ccelec sum q^2 / sqrt(2 pi sigma^2), out[1] = -ccelec pi Q^2 sigma^2 / V
hipError_t launch_corrections_finalize(const double* partials, int nblocks, double sigma,
                                       double bx, double by, double bz, double ccelec,
                                       double* out, hipStream_t s);
// sums n per-particle values of `ncomp` components laid out [ncomp][n]
hipError_t launch_sum_columns(const double* v, int n, int ncomp, double* partials,
                              hipStream_t s);
hipError_t launch_dipole(bool dbl, int X, int Y, int Z, double bx, double by,
                         double bz, const void* rho, double* partials,
                         hipStream_t s);
hipError_t launch_efield(bool dbl, int X, int Y, int Z, double bx, double by,
                         double bz, const void* phi, void* ex, void* ey,
                         void* ez, hipStream_t s);

// convolution: pass = 0 (x), 1 (y), 2 (z); taps on device (2R+1)
hipError_t launch_conv_pass(bool dbl, int pass, int X, int Y, int Z, int R,
                            const void* taps, const void* in, void* out,
                            hipStream_t s);

// charge statistics for the closed-form corrections: out[0] = sum q,
// out[1] = sum q^2 (FP64 partials -> finalize)
hipError_t launch_charge_moments(bool dbl, const void* xyzq, int n,
                                 double* partials, hipStream_t s);

}  // namespace gse
