// Internal host-side declarations shared by the C-ABI layer and the kernel
// translation units.  Nothing here is part of the public boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gse {

// Workgroups of the k-pass (one [ncomp] partial sum each).
constexpr int kKpassBlocksDefault = 512;

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

// Launch-shape knobs (process-wide; the autotuner sets them per workload).
struct Tuning {
  int bin_edge = 8;            // preferred bin edge in cells
  int brick_budget = 64 * 1024;  // LDS bytes allowed for an int64 brick
  int spread_block = 0;        // threads per spread workgroup (0: 512 FP32 / 256 FP64)
  int gather_block = 0;        // threads per gather workgroup (0: 512 FP32 / 256 FP64)
  int spread_stage = 16 * 1024;  // LDS bytes for staged particles beside the brick
  int gather_stage = 0;        // 0: 24 KiB for FP32 grids, 16 KiB for FP64
  int spread_lds32 = 1;        // FP32 grids: 32-bit accumulators with a per-bin scale
  int spread_cols = 0;         // column-owner spread in registers (else LDS-atomic brick)
  int spread_flush = 1;        // 1: bricks stored and summed by k_brick_reduce; 0: atomics
  int reduce_tile = 1;         // tiled brick reduce where the grid allows it
  int reduce_nt = 0;           // FP32 tiled reduce: 1 non-temporal brick loads
  int reduce_tile2 = 1;        // 64-bit bricks: tiled reduce with 16-byte pieces
  int bin_agg = 4;             // particles per workgroup / 256 of the aggregated binning (0: off)
  int fused_kz = 1;            // FFT solve as 2D x-y transforms around one fused z pass
  int zpass_cols = 8;          // spectrum columns per z-pass workgroup (4, 8, 16)
  int zpass4 = 1;              // four-step z pass for the column lengths it is built for
  int zpass4_cols = 0;         // spectrum columns per four-step workgroup (0 auto, 16, 32)
  int xy_plane = 0;            // x-y transforms of the fused solve by the plane kernels
  int ypass4 = 1;              // fused solve: 1D x rocFFT + four-step y pass where built
  int bin_coarse_block = 1024; // threads per super-bin binning workgroup (256, 512, 1024)
  int bin_order = 1;           // spread/gather workgroups take bins by descending count
  int gather_group = 32;       // lanes per charge in the FP32 13^3 gather walk (32, 16)
  int spread_mma = 0;          // FP32 spread on the matrix cores (int8 limbs, exact sums)
};
Tuning& tuning();
// LDS bytes for the gather's staged particles (the gather_stage knob, or its
// default): the FP32 brick (32 KB) leaves room for a bin's charges in one
// chunk; the FP64 brick (64 KB) keeps two workgroups per CU at 16 KiB
inline int gather_stage_bytes(bool dbl) {
  const int v = tuning().gather_stage;
  return v ? v : (dbl ? 16 * 1024 : 24 * 1024);
}

// Per-binning statistics, accumulated by the binning kernel and read by its
// consumers (spread, conversion, gather).  Two copies alternate between
// consecutive binnings; the consumers of one clear the other.
struct BinStats {
  unsigned long long qmax_bits;  // bits of max |q| (as a positive double)
  int invalid;                   // particles rejected (non-finite / out of range)
  int overflow;                  // particles that did not fit their bin
  int spill;                     // particles that did not fit their super-bin
};

// Particles binned into fixed-capacity bins: bin b owns slots
// [b*cap_bin, (b+1)*cap_bin) of sorted/perm; counts[b] may exceed cap_bin,
// the excess goes to the overflow list (indices into xyzq).
struct Binning {
  int n = 0;                 // particles in the call
  int cap_bin = 0;
  const void* xyzq = nullptr;  // the caller's particles
  int* counts = nullptr;       // [nbins] (current buffer)
  int* counts_other = nullptr; // [nbins] cleared by this binning's consumers
  BinStats* stats = nullptr;
  BinStats* stats_other = nullptr;
  int* ovf = nullptr;          // [n]
  int* perm = nullptr;         // [nbins*cap_bin]
  int* loc = nullptr;          // [nbins*cap_bin] cube origin in the bin: ox | oy << 8 | oz << 16
  void* sorted = nullptr;      // [nbins*cap_bin] 4-vectors of the compute type
  int images = 1;              // cube images per grid point (2n+1 > M wraps)
  double pref_bound = 0.0;     // bound of |q|-free contribution: pref * max weight
  void* bricks = nullptr;      // [nbins][Lz][Ly][Lx] LDS bricks (NULL: atomic flush)
  int* bshift = nullptr;       // [nbins] brick shift to 2^e, -1 for an empty bin
  // super-bin scratch of grids above 16384 bins (bin_sort_bytes): claim
  // counters, then the charges and their indices grouped by super-bin
  void* scan = nullptr;
  // [nbins] the bin each spread/gather workgroup position takes (NULL: bin =
  // position): the schedule from an earlier binning of the same shape (any
  // permutation of the bins gives the same results)
  int* order = nullptr;
  // [nbins] where the gather's spare workgroups write the schedule of this
  // binning's counts (NULL: none)
  int* order_next = nullptr;
};

// Fixed-point exponent e (scale 2^e) of one spread: every contribution
// |q pref w w w| stays below 2^51 and any grid point's sum below 2^62.
__host__ __device__ inline int fixed_exponent(unsigned long long qbits, int nvalid, int images,
                                              double pref_bound) {
  union {
    unsigned long long u;
    double d;
  } cv;
  cv.u = qbits;
  const double one = cv.d * pref_bound;
  if (!(one > 0.0)) return 60;
  const double all = one * (double)(nvalid > 0 ? nvalid : 1) * (double)images;
  int k1, k2;
  frexp(one, &k1);  // one < 2^k1
  frexp(all, &k2);
  int e = 51 - k1 < 62 - k2 ? 51 - k1 : 62 - k2;
  if (e < -60) e = -60;
  if (e > 200) e = 200;
  return e;
}

// Exponent of a 32-bit brick holding the sum of `cnt` contributions, each at
// most qmax * pref_bound: the sum stays below 2^30.
__host__ __device__ inline int fixed_exponent_bin(unsigned long long qbits, int cnt,
                                                  double pref_bound) {
  union {
    unsigned long long u;
    double d;
  } cv;
  cv.u = qbits;
  const double all = cv.d * pref_bound * (double)(cnt > 0 ? cnt : 1);
  if (!(all > 0.0)) return 30;
  int k;
  frexp(all, &k);  // all < 2^k
  return 30 - k;
}

// Device scratch of the allocation-free spread: one file-scope array per GPU
// (noalloc.hip).  noalloc_scratch returns the current device's copy, looked
// up once per device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t noalloc_scratch(void** out);

// ---- kernel launchers (defined in the .hip units) ------------------------
// ct_double: particle/grid compute type is double (else float)
// Clears counts[0..nbins) and the statistics of one buffer.
hipError_t launch_bin_clear(int* counts, int nbins, BinStats* stats, hipStream_t s);
// bytes of Binning::scan (super-bin scratch of grids above 16384 bins) for
// n particles of the compute type (dbl: double) in shape b (0: not needed);
// its first bytes are counters that must be zero when first used (the
// binning leaves them zero)
size_t bin_sort_bytes(int n, const BinShape& b, bool dbl);
hipError_t launch_bin(bool ct_double, const Geom& g, const BinShape& b, Binning& w,
                      hipStream_t s);

hipError_t launch_spread(bool ct_double, const Geom& g, const BinShape& b,
                         double sigma, bool normalize, Binning& w,
                         unsigned long long* grid64, hipStream_t s);

// Largest per-contribution weight factor of a spread (1 without normalize;
// with normalize each axis weight is at most sqrt(2 pi) s exp(h^2/8s^2)).
double spread_weight_bound(const Geom& g, double sigma, bool normalize);

hipError_t launch_fixed_to_grid(bool ct_double, const unsigned long long* grid64,
                                void* out, size_t count, const Binning& w,
                                bool clear, hipStream_t s);

// Brick-store flush: usable for this shape (LDS path, bounded brick overlap)?
bool brick_reduce_ok(const Geom& g, const BinShape& b);
bool spread_mma_ok(bool dbl, const Geom& g, const BinShape& b, bool normalize, bool acc32,
                   bool bricks);
// Which kernel converts a spread's fixed-point sums into the grid:
enum ReducePath : int {
  kReduceAtomic = 0,   // bricks added with 64-bit atomics, then k_fixed_to_grid
  kReduceGeneric = 1,  // k_brick_reduce, one thread per grid point
  kReduceTile4 = 2,    // k_brick_reduce_tile4 (32-bit bricks, 16-byte pieces)
  kReduceTile4NT = 3,  // the same with non-temporal brick loads
  kReduceTile2 = 4,    // k_brick_reduce_tile2 (64-bit bricks, two x points per thread)
  kReduceTile = 5,     // k_brick_reduce_tile (one x point per thread)
};
int brick_reduce_path(bool ct_double, bool acc32, const Geom& g, const BinShape& b);
// rho = 2^-e (sum of the bricks covering each point + the overflow int64 grid)
hipError_t launch_brick_reduce(bool ct_double, bool acc32, const Geom& g, const BinShape& b,
                               const Binning& w, unsigned long long* grid64, void* out,
                               hipStream_t s);

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
  // optional k-pass finalize run by one spare workgroup of the gather
  const double* fin_partials = nullptr;
  int fin_nblocks = kKpassBlocksDefault;  // partial sums to reduce
  int fin_ncomp = 0;
  double fin_pref = 0.0;
  double* fin_energy = nullptr;
  double* fin_virial = nullptr;
};

hipError_t launch_gather(bool ct_double, bool at_double, const Geom& g,
                         const BinShape& b, double sigma, double rcut,
                         int flags, const Binning& w, const void* phi,
                         const GatherOut& o, hipStream_t s);

// Number of extra workgroups the spread and gather launch for overflow
// particles (processed without LDS staging).
constexpr int kOverflowWG = 8;

// ---- batched independent systems -------------------------------------
// One descriptor per system, in device memory.  Particles of all systems are
// concatenated in the caller's system order (system s owns [pofs, pofs+n));
// grids, spectra and Green's tables are concatenated in the batch's internal
// order (systems with equal grid dimensions adjacent, so rocFFT runs them as
// one batched transform); bins and slots are numbered per system from bofs
// and slot0.
struct BatchSys {
  Geom g;
  BinShape b;
  int n;            // particles of the system
  int pofs;         // first particle in the concatenated arrays
  int bofs;         // first bin in the concatenated counts
  int cap_bin;      // slots per bin
  long long slot0;  // first slot in sorted / perm
  long long gofs;   // first grid point in grid64 / rho / phi
  long long kofs;   // first half-spectrum element in spectrum / table
  int images;
  int spread_cap;   // particles staged per chunk in the spread
  int gather_cap;   // particles staged per chunk in the gather
  int kblk0, nkblk; // k-pass chunks of this system
  long long brick0; // first brick element (brick-store flush)
  double inv2s2, pref, pref_bound, rc2, prefF, prefP, vcell;
};

struct BatchChunk {  // a contiguous range of one system's grid or spectrum
  int sys;
  int start;
  int count;
  int pad;
};

struct BatchDev {
  const BatchSys* sys = nullptr;
  int nsys = 0;
  const int* pofs = nullptr;  // [nsys+1] particle offsets (caller order)
  int ntot = 0;               // particles of all systems
  int* counts = nullptr;      // [total bins]
  int total_bins = 0;
  BinStats* stats = nullptr;  // [nsys]
  int* ovf = nullptr;         // [ntot] global particle indices
  int* novf = nullptr;        // [1]
  int* perm = nullptr;        // [slots] global particle index
  int* loc = nullptr;         // [slots] cube origin in the bin (ox | oy << 8 | oz << 16)
  void* sorted = nullptr;     // [slots] 4-vectors
  unsigned long long* grid64 = nullptr;
  void* bricks = nullptr;     // brick-store flush: [sum of nbins*vol] (NULL: atomic flush)
  int* bshift = nullptr;      // [total bins] brick shift, -1 for an empty bin
};

// clears counts, statistics and the overflow counter
hipError_t launch_batch_clear(const BatchDev& d, hipStream_t s);
hipError_t launch_batch_bin(bool dbl, const BatchDev& d, const void* xyzq, hipStream_t s);
// items: (system, bin) pairs of systems sharing 2n_z+1 = nz; lds: bytes
hipError_t launch_batch_spread(bool dbl, int nz, int block, size_t lds, int grid,
                               const BatchDev& d, const int2* items, int nitems, hipStream_t s);
// the overflow list of all systems (global atomics)
hipError_t launch_batch_spread_overflow(bool dbl, const BatchDev& d, const void* xyzq,
                                        hipStream_t s);
hipError_t launch_batch_convert(bool dbl, const BatchDev& d, const BatchChunk* chunks,
                                int nchunks, void* rho, hipStream_t s);
// brick-store flush: sums the bricks of the items' systems into rho (one
// workgroup per item; tile4: 32-bit bricks with rows of whole 16-byte pieces)
hipError_t launch_batch_reduce(bool dbl, bool acc32, bool tile4, const BatchDev& d,
                               const int2* items, int nitems, void* rho, hipStream_t s);
// the tiled reduce applies: bins of 8^3 on a grid of >= 3 whole bins per axis,
// bricks at most 3 bins wide
bool brick_tile_ok(const Geom& g, const BinShape& b);
hipError_t launch_batch_kpass(bool dbl, const BatchDev& d, const BatchChunk* chunks,
                              int nchunks, void* spec, const void* table, double* partials,
                              hipStream_t s);
hipError_t launch_batch_energy(const BatchDev& d, const double* partials, double ccelec,
                               double* energies, hipStream_t s);
hipError_t launch_batch_gather(bool ct_dbl, bool at_dbl, int nz, int block, size_t lds,
                               int grid, const BatchDev& d, const int2* items, int nitems,
                               const void* phi, const GatherOut& o, bool pair, hipStream_t s);
// FP32 gather of a cubic 13^3 support over 20 x 20 brick planes, or 15^3
// over 24 x 22: the two-charges-per-wave walk applies
bool gather_pair_ok(bool ct_dbl, bool at_dbl, const Geom& g, const BinShape& b);
hipError_t launch_batch_gather_overflow(bool ct_dbl, bool at_dbl, const BatchDev& d,
                                        const void* xyzq, const void* phi, const GatherOut& o,
                                        hipStream_t s);
// LDS bytes and particles staged per chunk of one system's spread / gather
size_t spread_lds_bytes(bool dbl, const Geom& g, const BinShape& b, int* cap);
size_t gather_lds_bytes(bool dbl, const Geom& g, const BinShape& b, int* cap);

// k-space
// zs: elements between z planes of the table (0: (X/2+1) Y, packed)
hipError_t launch_greens_table(bool dbl, int kind, int X, int Y, int Z,
                               double bx, double by, double bz, double sigma,
                               void* table, hipStream_t s, int zs = 0);
// plane stride of the fused solve's spectrum and G table: (X/2+1) Y rounded
// up to 32 elements, so every z plane starts on a 256-byte (FP32 complex) /
// 512-byte (FP64) boundary and the z pass reads whole cache lines
int fused_plane_stride(int X, int Y);

// In-place multiply of the half spectrum by the table, FP64 partial sums of
// the energy (and the 6 k-space virial moments when vir is set) per block.
// In-place multiply of the half spectrum by the table; FP64 partial sums of
// the energy (and, when vir is set, the six k-space virial moments) per
// block into partials[kKpassBlocks][vir ? 7 : 1].
constexpr int kKpassBlocks = kKpassBlocksDefault;
hipError_t launch_kpass(bool dbl, int kind, int X, int Y, int Z, double bx,
                        double by, double bz, double sigma, void* spec,
                        const void* table, bool vir, double* partials,
                        hipStream_t s);
// energy[0] = pref*E (may be null); virial[0..8] = W^k (ncomp 7 only) with
// pref = 0.5*ccelec*vcell, from nblocks partial sums.  The same reduction can
// run in a spare workgroup of the following gather (GatherOut::fin), saving
// a launch.
hipError_t launch_kpass_finalize(const double* partials, int nblocks, int ncomp, double pref,
                                 double* energy, double* virial, hipStream_t s);

// Fused z pass between 2D (x-y) transforms: z FFT, G(k) multiply with the
// k-pass's energy / virial partials, inverse z FFT, per column block.
constexpr int kZMaxStages = 8;
struct ZPlan {
  int Z, nst;
  int p[kZMaxStages];  // radices 2..8, product Z
  int C;               // spectrum columns per workgroup
  int four;            // 1: the four-step kernel built for this Z (k_zpass4)
};
// mixed-radix plan (radices 2..8, largest first) of a transform length;
// false when n has a prime factor above 7 or needs too many stages
bool radix_plan(int n, ZPlan* pl);
// false when Z has a prime factor above 7 or the columns do not fit LDS
bool zplan_make(int Z, bool dbl, ZPlan* pl);
size_t zpass_lds_bytes(const ZPlan& pl, bool dbl);
int zpass_blocks(int X, int Y, const ZPlan& pl);
// tw: device [Z] complex exp(-2 pi i m/Z) of the compute type; partials
// [zpass_blocks][vir ? 7 : 1]
// spec / table planes zs elements apart (0: (X/2+1) Y)
hipError_t launch_zpass(bool dbl, int kind, int X, int Y, int Z, double bx, double by, double bz,
                        double sigma, const ZPlan& pl, void* spec, const void* table,
                        const void* tw, bool vir, double* partials, hipStream_t s, int zs = 0);

// y transform of the half spectrum in registers (four-step, Y = 216 = 12 x 18):
// forward from packed planes (src, szs = (X/2+1) Y) into planes dzs apart,
// inverse the other way; tw: device [Y] complex exp(-2 pi i m/Y); also
// Y = 80 = 8 x 10
bool ypass4_shape(int Y, bool dbl);
hipError_t launch_ypass4(bool dbl, bool inv, int X, int Y, int Z, const void* src, int szs,
                         void* dst, int dzs, const void* tw, hipStream_t s);

// x-y plane transforms of the fused solve, one workgroup per z plane in LDS
// (in place of rocFFT's 2D r2c / c2r): rho [Z][Y][X] real <-> spec
// [Z][Y][X/2+1] complex, unnormalised.  xy_plane_ok: both lengths have radix
// plans and a plane fits the LDS budget.  twx / twy: device [X] / [Y]
// complex exp(-2 pi i m/n) of the compute type.
bool xy_plane_ok(int X, int Y, bool dbl);
hipError_t launch_xy_r2c(bool dbl, int X, int Y, int Z, const void* rho, void* spec,
                         const void* twx, const void* twy, hipStream_t s, int zs = 0);
hipError_t launch_xy_c2r(bool dbl, int X, int Y, int Z, const void* spec, void* phi,
                         const void* twx, const void* twy, hipStream_t s, int zs = 0);

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
// out[i] = f * in[i] (out may be in); out[i] += in[i]
hipError_t launch_grid_scale(bool dbl, void* out, const void* in, size_t n, double f,
                             hipStream_t s);
hipError_t launch_grid_add(bool dbl, void* out, const void* in, size_t n, hipStream_t s);
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
// sym: the taps are symmetric, k[R-i] == k[R+i] (FP32 sums pairs then)
hipError_t launch_conv_pass(bool dbl, int pass, int X, int Y, int Z, int R,
                            const void* taps, const void* in, void* out,
                            hipStream_t s, bool sym = false);

// charge statistics for the closed-form corrections: out[0] = sum q,
// out[1] = sum q^2 (FP64 partials -> finalize)
hipError_t launch_charge_moments(bool dbl, const void* xyzq, int n,
                                 double* partials, hipStream_t s);

// ---- SPME reciprocal comparator (ewald_pme.hip) -------------------------
// recip3: 1/Lx, 1/Ly, 1/Lz.  stats: max |q| and the rejected count of the
// call (cleared here), read by the spread and the int64 -> grid conversion.
constexpr int kPmeSumBlocks = 512;
hipError_t launch_pme_qmax(bool dbl, const void* xyzq, int n, const double* recip3,
                           BinStats* stats, hipStream_t s);
hipError_t launch_pme_spread(bool dbl, int order, const void* xyzq, int n, int X, int Y, int Z,
                             const double* recip3, const BinStats* stats,
                             unsigned long long* g64, hipStream_t s);
// prefac: device [X + Y + Z] B-spline moduli; partials [kPmeSumBlocks][7]
hipError_t launch_pme_scalar_sum(bool dbl, int X, int Y, int Z, const double* recip3,
                                 double kappa, const double* prefac, void* spec, bool want,
                                 double* partials, hipStream_t s);
hipError_t launch_pme_finalize(const double* partials, double ccelec, double* energy,
                               double* virial, hipStream_t s);
// out{x,y,z}[i] = sign * dE/dr_i (AT); invalid particles get 0
hipError_t launch_pme_gather(bool ct_dbl, bool at_dbl, int order, const void* xyzq, int n, int X,
                             int Y, int Z, const double* recip3, double ccelec, double sign,
                             const void* Qc, void* ox, void* oy, void* oz, hipStream_t s);

// ---- Ewald direct space (ewald_pme.hip) ----------------------------------
struct EwCells {
  int nc[3];      // cells per axis (1 when fewer than 3 cells of edge >= rcut fit)
  int ncell;
  double box[3];
};
struct EwPair {
  double rcut2, ccelec, inv_sqrt2sigma, inv_2sigmasq, two_sqrtpi;
};
// cell lists: counts/starts/fill [ncell], cell_of/tmp/sidx [n], sx [n] double4
hipError_t launch_ew_cells(const void* xyzq, int n, const EwCells& c, int* cell_of, int* counts,
                           int* starts, int* fill, int* tmp, int* sidx, void* sx,
                           hipStream_t s);
// xofs/xlist: symmetric exclusion lists (CSR over particles; xofs NULL: none)
hipError_t launch_ew_force(const EwCells& c, const EwPair& p, const int* starts,
                           const int* counts, const int* sidx, const void* sx, const int* xofs,
                           const int* xlist, double* gx, double* gy, double* gz,
                           double* partials, hipStream_t s);
hipError_t launch_ew_finalize(const double* partials, int ncell, const double* self2,
                              double* energy, hipStream_t s);

// ---- Ewald reciprocal sum (ewald_pme.hip) --------------------------------
struct ErArgs {
  int kcut, ncube;     // ncube = (2 kcut + 1)^3
  double ib[3];        // 1 / box
  double pref;         // 2 pi ccelec / V
  double half_sigmasq;
};
int ewald_recip_blocks(const ErArgs& a);
// sk: [ncube][3] (Re S, Im S, pek); partials: [ewald_recip_blocks]; forces
// -dE/dr (fx NULL: energy only), energy: device double
hipError_t launch_ewald_recip(const ErArgs& a, const void* xyzq, int n, double* sk,
                              double* partials, double* fx, double* fy, double* fz,
                              double* energy, hipStream_t s);

}  // namespace gse
