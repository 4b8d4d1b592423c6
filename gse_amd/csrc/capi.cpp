// Host runtime behind include/gse_hip.h: handles, device scratch, rocFFT
// plans and the stage sequencing of the GSE k-space step.  All device work
// is issued asynchronously on the handle's stream; only the calls documented
// as synchronising wait for it.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/gse_hip.h"
#include "gse_internal.h"

using namespace gse;

namespace gse {
extern Tuning g_tuning;
Tuning g_tuning;
Tuning& tuning() { return g_tuning; }
}  // namespace gse

namespace {

thread_local std::string g_err;

gse_status_t fail(gse_status_t s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? GSE_ERROR_OUT_OF_MEMORY : GSE_ERROR_HIP, \
                  "%s failed: %s", #expr, hipGetErrorString(e_));                      \
  } while (0)

#define GSECHK(expr)                   \
  do {                                 \
    gse_status_t s_ = (expr);          \
    if (s_ != GSE_SUCCESS) return s_;  \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  gse_status_t ensure(size_t need) {
    if (need <= bytes && p) return GSE_SUCCESS;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (need == 0) need = 256;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(GSE_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", need, hipGetErrorString(e));
    }
    bytes = need;
    return GSE_SUCCESS;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

bool valid_pair(gse_dtype_t at, gse_dtype_t ct) {
  if (at != GSE_FLOAT32 && at != GSE_FLOAT64) return false;
  if (ct != GSE_FLOAT32 && ct != GSE_FLOAT64) return false;
  return !(at == GSE_FLOAT32 && ct == GSE_FLOAT64);
}

size_t tsize(bool dbl) { return dbl ? 8 : 4; }

gse_status_t check_dims(int X, int Y, int Z) {
  if (X <= 0 || Y <= 0 || Z <= 0)
    return fail(GSE_ERROR_INVALID_VALUE, "grid sizes must be positive (got %d %d %d)", X, Y, Z);
  if ((double)X * Y * Z > 2.0e9)
    return fail(GSE_ERROR_INVALID_VALUE, "grid of %d x %d x %d points is too large", X, Y, Z);
  return GSE_SUCCESS;
}

gse_status_t check_box(double bx, double by, double bz) {
  if (!(bx > 0) || !(by > 0) || !(bz > 0) || !std::isfinite(bx) || !std::isfinite(by) ||
      !std::isfinite(bz))
    return fail(GSE_ERROR_INVALID_VALUE, "box sizes must be positive and finite");
  return GSE_SUCCESS;
}

gse_status_t check_sigma_rcut(double sigma, double rcut) {
  if (!(sigma > 0) || !std::isfinite(sigma))
    return fail(GSE_ERROR_INVALID_VALUE, "sigma must be positive and finite");
  if (!(rcut > 0) || !std::isfinite(rcut))
    return fail(GSE_ERROR_INVALID_VALUE, "rcut must be positive and finite");
  return GSE_SUCCESS;
}

// ---------------------------------------------------------------------------
// Particle engine: binning, fixed-point spread, gather
// ---------------------------------------------------------------------------
struct Engine {
  bool ct_dbl = true, at_dbl = true;
  int X = 0, Y = 0, Z = 0;
  DevBuf counts[2], stats[2], ovf, perm, loc, sorted, grid64, vir6, part, bricks, bshift,
      binscan, border[2];
  // schedule of the per-bin workgroups: border[ocur] holds one for order_nb
  // bins when order_ok; the gather writes the next into border[ocur ^ 1]
  int order_nb = 0, ocur = 0;
  bool order_ok = false;
  bool brick_flush = false;     // the last spread stored bricks (k_brick_reduce converts)
  int reduce_path = -1;         // ReducePath of the last conversion (-1: none yet)
  int spread_path = -1;         // last spread: 0 LDS integer atomics, 1 matrix cores
  int cur = 1;                  // buffer of the latest binning
  long long clean[2] = {0, 0};  // leading counts (and the stats) known zero
  bool grid64_clean = false;
  Binning bin;
  Geom geom{};
  BinShape shape{};

  void release() {
    for (DevBuf* b : {&counts[0], &counts[1], &stats[0], &stats[1], &ovf, &perm, &loc, &sorted,
                      &grid64, &vir6, &part, &bricks, &bshift, &binscan, &border[0],
                      &border[1]})
      b->release();
    order_nb = 0;
    order_ok = false;
  }

  // Bins the particles for a support of radius rcut and Gaussian sigma.
  gse_status_t bin_particles(int n, const void* xyzq, double bx, double by, double bz,
                             double sigma, double rcut, bool normalize, hipStream_t s) {
    if (n < 0) return fail(GSE_ERROR_INVALID_VALUE, "numCoord must be >= 0");
    if (n > 0 && !xyzq) return fail(GSE_ERROR_INVALID_VALUE, "xyzq is NULL");
    GSECHK(check_box(bx, by, bz));
    GSECHK(check_sigma_rcut(sigma, rcut));
    const Geom g = make_geom(X, Y, Z, bx, by, bz, rcut);
    geom = g;
    shape = choose_bins(g, 8, (size_t)g_tuning.brick_budget, g_tuning.bin_edge);
    const int nb = shape.nbins;
    int cap_bin;
    if (shape.direct) {
      cap_bin = n > 0 ? n : 1;
    } else {  // twice the mean occupancy plus a margin; excess goes to the overflow list
      const long long c = 2LL * n / nb + 32;
      cap_bin = (int)((c + 15) / 16 * 16);
    }
    const size_t slots = (size_t)nb * cap_bin;
    GSECHK(sorted.ensure(4 * tsize(ct_dbl) * slots));
    GSECHK(perm.ensure(sizeof(int) * slots));
    GSECHK(loc.ensure(sizeof(int) * slots));
    GSECHK(ovf.ensure(sizeof(int) * (size_t)(n > 0 ? n : 1)));
    GSECHK(part.ensure(sizeof(double) * kRedBlocks * 8));
    const size_t sb = bin_sort_bytes(n, shape, ct_dbl);
    if (sb) {
      const void* before = binscan.p;
      GSECHK(binscan.ensure(sb));
      // the super-bin counters start at zero; the binning leaves them zero
      if (binscan.p != before) HIPCHK(hipMemset(binscan.p, 0, binscan.bytes));
    }
    for (int k = 0; k < 2; k++) {
      if (counts[k].bytes < sizeof(int) * (size_t)nb || !counts[k].p) {
        GSECHK(counts[k].ensure(sizeof(int) * (size_t)nb));
        clean[k] = 0;
      }
      if (!stats[k].p) {
        GSECHK(stats[k].ensure(sizeof(BinStats)));
        clean[k] = 0;
      }
    }
    const size_t npts = (size_t)X * Y * Z;
    if (grid64.bytes < npts * 8 || !grid64.p) {
      GSECHK(grid64.ensure(npts * 8));
      grid64_clean = false;
    }
    if (!grid64_clean) {
      HIPCHK(launch_fill_u64(grid64.as<unsigned long long>(), npts, s));
      grid64_clean = true;
    }
    const int c = cur ^ 1;
    if (clean[c] < nb) HIPCHK(launch_bin_clear(counts[c].as<int>(), nb, stats[c].as<BinStats>(), s));
    clean[c] = 0;
    cur = c;
    bin.n = n;
    bin.cap_bin = cap_bin;
    bin.xyzq = xyzq;
    bin.counts = counts[c].as<int>();
    bin.counts_other = counts[c ^ 1].as<int>();
    bin.stats = stats[c].as<BinStats>();
    bin.stats_other = stats[c ^ 1].as<BinStats>();
    bin.ovf = ovf.as<int>();
    bin.perm = perm.as<int>();
    bin.loc = loc.as<int>();
    bin.sorted = sorted.p;
    bin.scan = sb ? binscan.p : nullptr;
    bin.images = ((2 * g.nx + 1 + X - 1) / X) * ((2 * g.ny + 1 + Y - 1) / Y) *
                 ((2 * g.nz + 1 + Z - 1) / Z);
    bin.pref_bound = std::pow(2.0 * M_PI * sigma * sigma, -1.5) *
                     spread_weight_bound(g, sigma, normalize);
    HIPCHK(launch_bin(ct_dbl, g, shape, bin, s));
    // the spread's and the gather's workgroups take bins by descending count
    // of the last binning of this shape; the gather's spare workgroups sort
    // this binning's counts for the next
    bin.order = nullptr;
    bin.order_next = nullptr;
    if (!shape.direct && n > 0 && g_tuning.bin_order) {
      if (order_nb != nb) {
        GSECHK(border[0].ensure(sizeof(int) * (size_t)nb));
        GSECHK(border[1].ensure(sizeof(int) * (size_t)nb));
        order_nb = nb;
        order_ok = false;
      }
      bin.order = order_ok ? border[ocur].as<int>() : nullptr;
      bin.order_next = border[ocur ^ 1].as<int>();
    }
    return GSE_SUCCESS;
  }

  // the consumer kernels clear the other buffer
  void consumed() { clean[cur ^ 1] = shape.nbins; }

  bool acc32() const { return !ct_dbl && g_tuning.spread_lds32; }

  gse_status_t spread_binned(double sigma, bool normalize, hipStream_t s) {
    brick_flush = g_tuning.spread_flush && !g_tuning.spread_cols && brick_reduce_ok(geom, shape);
    bin.bricks = nullptr;
    bin.bshift = nullptr;
    if (brick_flush) {
      const size_t vol = (size_t)shape.Lx * shape.Ly * shape.Lz;
      GSECHK(bricks.ensure((acc32() ? 4 : 8) * vol * (size_t)shape.nbins));
      GSECHK(bshift.ensure(sizeof(int) * (size_t)shape.nbins));
      bin.bricks = bricks.p;
      bin.bshift = bshift.as<int>();
    }
    spread_path = !g_tuning.spread_cols && g_tuning.spread_mma &&
                          spread_mma_ok(ct_dbl, geom, shape, normalize, acc32(), brick_flush)
                      ? 1
                      : 0;
    HIPCHK(launch_spread(ct_dbl, geom, shape, sigma, normalize, bin,
                         grid64.as<unsigned long long>(), s));
    consumed();
    return GSE_SUCCESS;
  }

  gse_status_t convert(void* rho, hipStream_t s) {
    if (brick_flush) {
      reduce_path = brick_reduce_path(ct_dbl, acc32(), geom, shape);
      HIPCHK(launch_brick_reduce(ct_dbl, acc32(), geom, shape, bin,
                                 grid64.as<unsigned long long>(), rho, s));
    } else {
      reduce_path = kReduceAtomic;
      HIPCHK(launch_fixed_to_grid(ct_dbl, grid64.as<unsigned long long>(), rho,
                                  (size_t)X * Y * Z, bin, true, s));
    }
    return GSE_SUCCESS;
  }

  // [Bx, By, Bz, nbins, Lx, Ly, Lz, direct, reduce path, slots per bin,
  //  k-space form, spread kernel]
  void layout(int* info, int kform = -1) const {
    const int v[GSE_LAYOUT_FIELDS] = {shape.Bx, shape.By, shape.Bz, shape.nbins, shape.Lx,
                                      shape.Ly, shape.Lz, shape.direct, reduce_path, bin.cap_bin,
                                      kform, spread_path};
    std::memcpy(info, v, sizeof(v));
  }

  gse_status_t gather_binned(double sigma, double rcut, int flags, const void* phi,
                             GatherOut& o, double* virial9, hipStream_t s) {
    const int n = bin.n;
    if (flags & GF_VIRIAL) {
      GSECHK(vir6.ensure(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1)));
      o.vir6 = vir6.as<double>();
      o.vir_stride = n > 0 ? n : 1;
      if (n > 0)  // invalid particles leave their slots untouched
        HIPCHK(launch_fill(true, vir6.p, 6 * (size_t)n, s));
    }
    HIPCHK(launch_gather(ct_dbl, at_dbl, geom, shape, sigma, rcut, flags, bin, phi, o, s));
    consumed();
    if (bin.order_next) {  // the schedule of these counts is ready
      ocur ^= 1;
      order_ok = true;
      bin.order = border[ocur].as<int>();
      bin.order_next = border[ocur ^ 1].as<int>();
    }
    if ((flags & GF_VIRIAL) && virial9 && n > 0) {
      HIPCHK(launch_sum_columns(vir6.as<double>(), n, 6, part.as<double>(), s));
      HIPCHK(launch_add_sym6(part.as<double>(), kRedBlocks, virial9, s));
    }
    return GSE_SUCCESS;
  }

  gse_status_t invalid_count(hipStream_t s, int* out) {
    if (!stats[cur].p) {
      *out = 0;
      return GSE_SUCCESS;
    }
    BinStats h;
    HIPCHK(hipMemcpyAsync(&h, stats[cur].p, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *out = h.invalid;
    return GSE_SUCCESS;
  }
};

// ---------------------------------------------------------------------------
// rocFFT: 3D out-of-place r2c / c2r, lengths fastest first {X, Y, Z}
// ---------------------------------------------------------------------------
std::once_flag g_rocfft_once;

// Process-wide cache of rocFFT plan pairs keyed by (precision, X, Y, Z,
// batch).  Handles of equal shape share one plan pair (reference counted);
// each handle keeps its own execution info (stream) and work buffer.  A plan
// pair is destroyed when its last handle releases it.
struct PlanPair {
  rocfft_plan fwd = nullptr, bwd = nullptr;
  size_t work_f = 0, work_b = 0;
  int refs = 0;
};
struct PlanKey {
  int dbl, X, Y, Z, batch;
  int dim = 3;  // 3: 3D transforms; 2: 2D x-y transforms of the Z planes
  int cdist = 0;  // 2D: elements between the half spectrum's z planes (0: (X/2+1) Y)
  bool operator<(const PlanKey& o) const {
    return std::tie(dbl, X, Y, Z, batch, dim, cdist) <
           std::tie(o.dbl, o.X, o.Y, o.Z, o.batch, o.dim, o.cdist);
  }
};
std::mutex g_plan_mu;
std::map<PlanKey, PlanPair> g_plans;

gse_status_t create_plan_pair(const PlanKey& k, PlanPair& pp) {
  const int X = k.X, Y = k.Y, Z = k.Z;
  const int batch = k.dim == 2 ? Z : k.dim == 1 ? Y * Z : k.batch;
  const size_t lengths[3] = {(size_t)X, (size_t)Y, (size_t)Z};
  const rocfft_precision prec = k.dbl ? rocfft_precision_double : rocfft_precision_single;
  rocfft_plan_description df = nullptr, db = nullptr;
  if (k.dim == 2) {
    // the Z x-y planes of a [z][y][x] grid into the [z][y][x <= X/2] half
    // spectrum: the layout the 3D r2c leaves, before its z transform, with
    // its planes k.cdist elements apart when that is set
    const size_t Xh = (size_t)X / 2 + 1;
    const size_t rs[2] = {1, (size_t)X};
    const size_t cs[2] = {1, Xh};
    const size_t rd = (size_t)X * Y, cd = k.cdist > 0 ? (size_t)k.cdist : Xh * Y;
    if (rocfft_plan_description_create(&df) != rocfft_status_success ||
        rocfft_plan_description_create(&db) != rocfft_status_success ||
        rocfft_plan_description_set_data_layout(df, rocfft_array_type_real,
                                                rocfft_array_type_hermitian_interleaved,
                                                nullptr, nullptr, 2, rs, rd, 2, cs, cd) !=
            rocfft_status_success ||
        rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved,
                                                rocfft_array_type_real, nullptr, nullptr, 2, cs,
                                                cd, 2, rs, rd) != rocfft_status_success) {
      if (df) rocfft_plan_description_destroy(df);
      if (db) rocfft_plan_description_destroy(db);
      return fail(GSE_ERROR_FFT, "rocfft 2D plan description (%dx%d x %d) failed", X, Y, Z);
    }
  } else if (k.dim == 1) {
    // the Y Z rows of a [z][y][x] grid into packed [z][y][x <= X/2] rows
    const size_t Xh = (size_t)X / 2 + 1;
    const size_t one[1] = {1};
    if (rocfft_plan_description_create(&df) != rocfft_status_success ||
        rocfft_plan_description_create(&db) != rocfft_status_success ||
        rocfft_plan_description_set_data_layout(df, rocfft_array_type_real,
                                                rocfft_array_type_hermitian_interleaved,
                                                nullptr, nullptr, 1, one, (size_t)X, 1, one,
                                                Xh) != rocfft_status_success ||
        rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved,
                                                rocfft_array_type_real, nullptr, nullptr, 1, one,
                                                Xh, 1, one, (size_t)X) != rocfft_status_success) {
      if (df) rocfft_plan_description_destroy(df);
      if (db) rocfft_plan_description_destroy(db);
      return fail(GSE_ERROR_FFT, "rocfft 1D plan description (%d x %d) failed", X, Y * Z);
    }
  } else if (batch > 1) {
    const size_t Xh = (size_t)X / 2 + 1;
    const size_t rs[3] = {1, (size_t)X, (size_t)X * Y};
    const size_t cs[3] = {1, Xh, Xh * Y};
    const size_t rd = (size_t)X * Y * Z, cd = Xh * Y * Z;
    if (rocfft_plan_description_create(&df) != rocfft_status_success ||
        rocfft_plan_description_create(&db) != rocfft_status_success ||
        rocfft_plan_description_set_data_layout(df, rocfft_array_type_real,
                                                rocfft_array_type_hermitian_interleaved,
                                                nullptr, nullptr, 3, rs, rd, 3, cs, cd) !=
            rocfft_status_success ||
        rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved,
                                                rocfft_array_type_real, nullptr, nullptr, 3,
                                                cs, cd, 3, rs, rd) != rocfft_status_success) {
      if (df) rocfft_plan_description_destroy(df);
      if (db) rocfft_plan_description_destroy(db);
      return fail(GSE_ERROR_FFT, "rocfft plan description (%dx%dx%d x %d) failed", X, Y, Z,
                  batch);
    }
  }
  const size_t ndim = (size_t)k.dim;
  const rocfft_status sf =
      rocfft_plan_create(&pp.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                         prec, ndim, lengths, (size_t)batch, df);
  const rocfft_status sb =
      rocfft_plan_create(&pp.bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                         prec, ndim, lengths, (size_t)batch, db);
  if (df) rocfft_plan_description_destroy(df);
  if (db) rocfft_plan_description_destroy(db);
  if (sf != rocfft_status_success || sb != rocfft_status_success) {
    if (sf == rocfft_status_success) rocfft_plan_destroy(pp.fwd);
    if (sb == rocfft_status_success) rocfft_plan_destroy(pp.bwd);
    pp.fwd = pp.bwd = nullptr;
    return fail(GSE_ERROR_FFT, "rocfft_plan_create (%s %dx%dx%d x %d) failed",
                sf != rocfft_status_success ? "r2c" : "c2r", X, Y, Z, batch);
  }
  rocfft_plan_get_work_buffer_size(pp.fwd, &pp.work_f);
  rocfft_plan_get_work_buffer_size(pp.bwd, &pp.work_b);
  return GSE_SUCCESS;
}

gse_status_t acquire_plans(const PlanKey& k, PlanPair* out) {
  std::call_once(g_rocfft_once, [] { rocfft_setup(); });
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(k);
  if (it == g_plans.end()) {
    PlanPair pp;
    GSECHK(create_plan_pair(k, pp));
    it = g_plans.emplace(k, pp).first;
  }
  it->second.refs++;
  *out = it->second;
  return GSE_SUCCESS;
}

void release_plans(const PlanKey& k) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(k);
  if (it == g_plans.end()) return;
  if (--it->second.refs == 0) {
    rocfft_plan_destroy(it->second.fwd);
    rocfft_plan_destroy(it->second.bwd);
    g_plans.erase(it);
  }
}

struct FFT3 {
  bool dbl = true;
  int X = 0, Y = 0, Z = 0;
  PlanKey key{};
  rocfft_plan fwd = nullptr, bwd = nullptr;  // shared, owned by the cache
  rocfft_execution_info info = nullptr;
  DevBuf work;
  size_t work_f = 0, work_b = 0;

  // batch > 1: that many transforms of packed, back-to-back grids (distance
  // X*Y*Z reals) and half spectra (distance (X/2+1)*Y*Z complex).
  // dim = 2: the 2D x-y transforms of the Z planes only.
  // cdist (dim 2): elements between the half spectrum's z planes
  gse_status_t init(bool d, int x, int y, int z, int batch = 1, int dim = 3, int cdist = 0) {
    dbl = d;
    X = x;
    Y = y;
    Z = z;
    key = PlanKey{d ? 1 : 0, x, y, z, batch, dim, cdist};
    PlanPair pp;
    GSECHK(acquire_plans(key, &pp));
    fwd = pp.fwd;
    bwd = pp.bwd;
    work_f = pp.work_f;
    work_b = pp.work_b;
    const size_t w = work_f > work_b ? work_f : work_b;
    if (w) GSECHK(work.ensure(w));
    if (rocfft_execution_info_create(&info) != rocfft_status_success)
      return fail(GSE_ERROR_FFT, "rocfft_execution_info_create failed");
    return GSE_SUCCESS;
  }

  void release() {
    if (fwd) release_plans(key);
    if (info) rocfft_execution_info_destroy(info);
    fwd = bwd = nullptr;
    info = nullptr;
    work.release();
  }

  gse_status_t exec(bool forward, const void* in, void* out, hipStream_t s) {
    rocfft_plan p = forward ? fwd : bwd;
    const size_t w = forward ? work_f : work_b;
    if (rocfft_execution_info_set_stream(info, s) != rocfft_status_success)
      return fail(GSE_ERROR_FFT, "rocfft_execution_info_set_stream failed");
    if (w && rocfft_execution_info_set_work_buffer(info, work.p, w) != rocfft_status_success)
      return fail(GSE_ERROR_FFT, "rocfft_execution_info_set_work_buffer failed");
    void* ib[1] = {const_cast<void*>(in)};
    void* ob[1] = {out};
    if (rocfft_execute(p, ib, ob, info) != rocfft_status_success)
      return fail(GSE_ERROR_FFT, "rocfft_execute failed");
    return GSE_SUCCESS;
  }
};

}  // namespace
// rocFFT pair of one SPME handle (the plans come from the shared cache)
struct KSolverFFTOnly {
  FFT3 fft;
};
struct gse_gsek_s;
static gse_status_t stage_mark(gse_gsek_s* h);
namespace {

// k-space solve state: spectrum, Green's table, partial sums
struct KSolver {
  bool dbl = true;
  int X = 0, Y = 0, Z = 0;
  FFT3 fft;
  // fused form: 2D x-y transforms around one z pass (ZPlan), when Z allows it
  bool fused = false;
  FFT3 fft2;
  ZPlan zp{};
  DevBuf tw;
  // x-y transforms of the fused form by the plane kernels (else rocFFT 2D)
  bool plane = false;
  DevBuf twx, twy;
  // else, where built: 1D x rocFFT into packed spec2, then the four-step y
  // pass into the padded spectrum (and back)
  bool ypass = false;
  DevBuf spec2;
  int zblocks = 0;
  int zs = 0;  // elements between the spectrum's (and G table's) z planes
  DevBuf spec, table, part;
  int tab_kind = -1;
  double tab_sigma = NAN, tab_bx = NAN, tab_by = NAN, tab_bz = NAN;

  size_t nk() const { return (size_t)(X / 2 + 1) * Y * Z; }
  // 0: 3D transforms + k-pass, 1: fused Stockham z pass, 2: fused four-step;
  // +2 when the plane kernels do the x-y transforms, +4 with the 1D x
  // rocFFT transforms and the four-step y pass
  int form() const { return fused ? (zp.four ? 2 : 1) + (plane ? 2 : 0) + (ypass ? 4 : 0) : 0; }

  // table exp(-2 pi i m/n), m < n, in FP64 rounded to the compute type
  gse_status_t upload_twiddles(DevBuf& b, int n) {
    std::vector<double> w(2 * (size_t)n);
    for (int m = 0; m < n; m++) {
      w[2 * m] = std::cos(2.0 * M_PI * m / n);
      w[2 * m + 1] = -std::sin(2.0 * M_PI * m / n);
    }
    GSECHK(b.ensure(2 * tsize(dbl) * (size_t)n));
    if (dbl) {
      HIPCHK(hipMemcpy(b.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    } else {
      std::vector<float> wf(w.begin(), w.end());
      HIPCHK(hipMemcpy(b.p, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice));
    }
    return GSE_SUCCESS;
  }

  gse_status_t init(bool d, int x, int y, int z) {
    dbl = d;
    X = x;
    Y = y;
    Z = z;
    GSECHK(fft.init(d, x, y, z));
    size_t nparts = kKpassBlocks;
    fused = g_tuning.fused_kz && zplan_make(z, d, &zp);
    // the fused form's planes start on whole cache lines (the z pass reads
    // one piece of every plane per column group); the 3D form's are packed
    zs = fused ? fused_plane_stride(x, y) : (x / 2 + 1) * y;
    GSECHK(spec.ensure(2 * tsize(d) * (size_t)zs * z));
    GSECHK(table.ensure(tsize(d) * (size_t)zs * z));
    if (fused) {
      plane = g_tuning.xy_plane && xy_plane_ok(x, y, d);
      ypass = !plane && g_tuning.ypass4 && ypass4_shape(y, d);
      if (plane) {
        GSECHK(upload_twiddles(twx, x));
        GSECHK(upload_twiddles(twy, y));
      } else if (ypass) {
        GSECHK(fft2.init(d, x, y, z, 1, 1));
        GSECHK(upload_twiddles(twy, y));
        GSECHK(spec2.ensure(2 * tsize(d) * nk()));
      } else {
        GSECHK(fft2.init(d, x, y, z, 1, 2, zs));
      }
      zblocks = zpass_blocks(x, y, zp);
      nparts = std::max<size_t>(nparts, (size_t)zblocks);
      GSECHK(upload_twiddles(tw, z));
    }
    GSECHK(part.ensure(sizeof(double) * nparts * 7));
    return GSE_SUCCESS;
  }
  void release() {
    fft.release();
    fft2.release();
    tw.release();
    twx.release();
    twy.release();
    spec2.release();
    spec.release();
    table.release();
    part.release();
  }
  gse_status_t ensure_table(int kind, double sigma, double bx, double by, double bz,
                            hipStream_t s) {
    if (kind == tab_kind && sigma == tab_sigma && bx == tab_bx && by == tab_by && bz == tab_bz)
      return GSE_SUCCESS;
    HIPCHK(launch_greens_table(dbl, kind, X, Y, Z, bx, by, bz, sigma, table.p, s, zs));
    tab_kind = kind;
    tab_sigma = sigma;
    tab_bx = bx;
    tab_by = by;
    tab_bz = bz;
    return GSE_SUCCESS;
  }
  // phi = c2r(G * r2c(rho)); energy/virial (device) when requested
  // energy / virial: device outputs (may be null); the k-space virial is
  // written into virial[0..8].  With `defer` set, the final reduction is not
  // launched but described in *defer (to run in the gather's spare workgroup).
  gse_status_t solve(int kind, double sigma, double bx, double by, double bz, const void* rho,
                     void* phi, double* energy, double* virial, double ccelec, hipStream_t s,
                     gse_gsek_s* marker = nullptr, GatherOut* defer = nullptr) {
    GSECHK(ensure_table(kind, sigma, bx, by, bz, s));
    FFT3& f = fused ? fft2 : fft;
    const int packed = (X / 2 + 1) * Y;
    if (plane) {
      HIPCHK(launch_xy_r2c(dbl, X, Y, Z, rho, spec.p, twx.p, twy.p, s, zs));
    } else if (ypass) {
      GSECHK(f.exec(true, rho, spec2.p, s));
      HIPCHK(launch_ypass4(dbl, false, X, Y, Z, spec2.p, packed, spec.p, zs, twy.p, s));
    } else {
      GSECHK(f.exec(true, rho, spec.p, s));
    }
    if (marker) GSECHK(stage_mark(marker));
    int nblocks = kKpassBlocks;
    if (fused) {
      HIPCHK(launch_zpass(dbl, kind, X, Y, Z, bx, by, bz, sigma, zp, spec.p, table.p, tw.p,
                          virial != nullptr, part.as<double>(), s, zs));
      nblocks = zblocks;
    } else {
      HIPCHK(launch_kpass(dbl, kind, X, Y, Z, bx, by, bz, sigma, spec.p, table.p,
                          virial != nullptr, part.as<double>(), s));
    }
    if (energy || virial) {
      const double vcell = (bx / X) * (by / Y) * (bz / Z);
      const int ncomp = virial ? 7 : 1;
      if (defer) {
        defer->fin_partials = part.as<double>();
        defer->fin_nblocks = nblocks;
        defer->fin_ncomp = ncomp;
        defer->fin_pref = 0.5 * ccelec * vcell;
        defer->fin_energy = energy;
        defer->fin_virial = virial;
      } else {
        HIPCHK(launch_kpass_finalize(part.as<double>(), nblocks, ncomp, 0.5 * ccelec * vcell,
                                     energy, virial, s));
      }
    }
    if (marker) GSECHK(stage_mark(marker));
    if (plane) {
      HIPCHK(launch_xy_c2r(dbl, X, Y, Z, spec.p, phi, twx.p, twy.p, s, zs));
    } else if (ypass) {
      HIPCHK(launch_ypass4(dbl, true, X, Y, Z, spec.p, zs, spec2.p, packed, twy.p, s));
      GSECHK(f.exec(false, spec2.p, phi, s));
    } else {
      GSECHK(f.exec(false, spec.p, phi, s));
    }
    if (marker) GSECHK(stage_mark(marker));
    return GSE_SUCCESS;
  }
};

}  // namespace

// ===========================================================================
// handles
// ===========================================================================
struct gse_gauss_charge_s {
  Engine eng;
  hipStream_t stream = nullptr;
  DevBuf tmpA, tmpB, taps, part, dev3;
};

struct gse_convolution_s {
  bool dbl = true;
  int X = 0, Y = 0, Z = 0;
  int R[3] = {-1, -1, -1};
  bool sym[3] = {false, false, false};  // taps symmetric about the centre
  DevBuf taps[3], tmpA, tmpB;
  hipStream_t stream = nullptr;
};

struct gse_fft_solver_s {
  KSolver ks;
  int kind = 0;
  hipStream_t stream = nullptr;
};

struct gse_gsek_s {
  bool at_dbl = true, ct_dbl = true;
  int X = 0, Y = 0, Z = 0;
  double sigma = 0, kappa = 0, ls = 0, ls1 = 0, ccelec = 332.0716;
  double bx = 0, by = 0, bz = 0;
  double sigma1 = 0, sigma2 = 0;
  Engine eng;
  KSolver ks;
  DevBuf rhoS, phiM, rhoR, taps[3], tmp, tmp2, dout, part;
  int tapsR[3] = {0, 0, 0};
  double taps_key[5] = {NAN, NAN, NAN, NAN, NAN};  // sigma2, rcut2, box: taps are current
  hipStream_t stream = nullptr;
  bool profiling = false;
  std::vector<hipEvent_t> events;  // (GSE_NUM_STAGES + 1) per recorded step
  size_t ev_used = 0;
  double rcut1() const { return ls1 * std::sqrt(2.0) * sigma1; }
  double rcut2() const { return ls1 * std::sqrt(2.0) * sigma2; }
};

// Batched independent systems: one layout of all systems' bins, grids and
// spectra, one launch per stage (per bin-shape group for spread and gather,
// per grid size for the FFTs), optionally replayed from a HIP graph.
struct BatchGroup {  // systems sharing 2nz+1: one spread and one gather launch
  int nz = 0;
  int item0 = 0, nitems = 0;
  size_t lds_spread = 0, lds_gather = 0;
  int grid = 0;
  bool tile4 = true;  // every brick row is whole 16-byte pieces (32-bit bricks)
  bool pair = true;   // every system takes the two-charges-per-wave gather
};

struct BatchFFT {  // systems with equal grid dimensions, adjacent in the layout
  FFT3 fft;
  long long gofs = 0, kofs = 0;
};

struct gse_batch_s {
  bool at_dbl = false, ct_dbl = false;
  int nsys = 0;
  double sigma = 0, kappa = 0, ls = 0, ls1 = 0, ccelec = 332.0716;
  double sigma1 = 0, sigma2 = 0, rc1 = 0;
  std::vector<int> dims, counts;
  std::vector<double> boxes;
  std::vector<int> order;  // internal order of systems (equal grids adjacent)
  int ntot = 0;
  // launch shape (set by gse_batch_set_config / the autotuner)
  int bin_edge = 8, spread_block = 0, gather_block = 0;
  // layout
  std::vector<BatchSys> sys;
  std::vector<BatchGroup> groups;
  std::vector<BatchFFT*> ffts;
  std::vector<int2> items;
  std::vector<BatchChunk> cchunks, kchunks;
  long long total_points = 0, total_k = 0, total_slots = 0;
  int total_bins = 0;
  DevBuf d_sys, d_pofs, d_items, d_cchunks, d_kchunks, counts_b, stats_b, ovf_b, novf_b, perm_b,
      loc_b, sorted_b, grid64_b, rho_b, phi_b, spec_b, table_b, part_b, bricks_b, bshift_b;
  bool brick_flush = false;  // bricks stored and summed by the tiled reduce
  BatchDev dev;
  hipStream_t stream = nullptr;
  // graph of one step
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_ptrs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

  void release_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
  }
  void release_ffts() {
    for (BatchFFT* f : ffts) {
      f->fft.release();
      delete f;
    }
    ffts.clear();
  }
  void release() {
    release_graph();
    release_ffts();
    for (DevBuf* b : {&d_sys, &d_pofs, &d_items, &d_cchunks, &d_kchunks, &counts_b, &stats_b,
                      &ovf_b, &novf_b, &perm_b, &loc_b, &sorted_b, &grid64_b, &rho_b, &phi_b,
                      &spec_b, &table_b, &part_b, &bricks_b, &bshift_b})
      b->release();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    cap_stream = nullptr;
  }
};

// SPME reciprocal comparator (cuda/CudaPMERecip.h:36-182, BOX FFT, one GPU)
struct gse_pme_recip_s {
  bool at_dbl = true, ct_dbl = true;
  int X = 0, Y = 0, Z = 0, order = 4;
  double recip3[3] = {0, 0, 0};  // of the last spread (gather and scalar_sum reuse it)
  int n_last = 0;
  KSolverFFTOnly* fft = nullptr;
  DevBuf grid64, grid, spec, prefac, partials, stats, tmp2;
  hipStream_t stream = nullptr;
};

// Ewald reciprocal sum (CpuEwaldRecip.h:6-40)
struct gse_ewald_recip_s {
  double sigma = 0, ccelec = 332.0716, box[3] = {0, 0, 0};
  int kcut = 0;
  DevBuf sk, partials;
  hipStream_t stream = nullptr;
};

// GSE-r (CpuGSEr.h): a GSE-k handle's engine, solver and grids (core->rhoS
// holds the sigma2 grid, core->phiM the potential) plus the sigma1 grid rho
struct gse_gser_s {
  gse_gsek_s* core = nullptr;
  DevBuf rho;
};

// GSE-u (CpuGSEu.h): the core's rhoS is the sigma1 grid; N convolved terms
struct gse_gseu_s {
  gse_gsek_s* core = nullptr;
  int N = 0;
  double b = 0.0;
  std::vector<double> sig, fac;        // sigma_j and (2 pi sigma_j^2)^(3/2)/sigma_j
  struct Taps3 {
    DevBuf t[3];
    int R[3];
  };
  std::vector<Taps3*> taps;            // the three tap arrays and radii of each term
  std::vector<DevBuf*> terms;          // rhoG_j
};

// Ewald direct space + self energy (CpuEwaldDirect.h:10-44)
struct gse_ewald_direct_s {
  double sigma = 0, rcut = 0, ccelec = 332.0716, box[3] = {0, 0, 0};
  DevBuf xofs, xlist, cell_of, counts, starts, fill, tmp, sidx, sx, partials, self2, mom;
  bool have_excl = false;
  int n_excl_atoms = 0;
  hipStream_t stream = nullptr;
};

namespace {

// separable sigma taps of spreadChargeOnGrid (CpuGaussCharge.cpp:406-433)
gse_status_t make_ongrid_taps(bool dbl, int X, int Y, int Z, double sigma, double rcut,
                              double bx, double by, double bz, DevBuf (&taps)[3], int (&R)[3]) {
  const double h[3] = {bx / X, by / Y, bz / Z};
  const double inv_2sigmasq = 1.0 / (2.0 * sigma * sigma);
  const double pref = std::pow(2.0 * M_PI * sigma * sigma, -1.5) * h[0] * h[1] * h[2];
  for (int a = 0; a < 3; a++) {
    R[a] = (int)std::ceil(rcut / h[a]);
    std::vector<double> kd(2 * R[a] + 1);
    const double hsq = h[a] * h[a];
    for (int ii = -R[a]; ii <= R[a]; ii++) {
      double v = std::exp(-hsq * (double)(ii * ii) * inv_2sigmasq);
      kd[ii + R[a]] = a == 2 ? pref * v : v;
    }
    GSECHK(taps[a].ensure(tsize(dbl) * kd.size()));
    if (dbl) {
      HIPCHK(hipMemcpy(taps[a].p, kd.data(), 8 * kd.size(), hipMemcpyHostToDevice));
    } else {
      std::vector<float> kf(kd.begin(), kd.end());
      HIPCHK(hipMemcpy(taps[a].p, kf.data(), 4 * kf.size(), hipMemcpyHostToDevice));
    }
  }
  return GSE_SUCCESS;
}

// on-grid sigma2 taps of a GSE-k handle, rebuilt only when sigma2, rcut2 or
// the box change (the upload is synchronous)
gse_status_t gsek_taps(gse_gsek_s* h) {
  const double key[5] = {h->sigma2, h->rcut2(), h->bx, h->by, h->bz};
  if (std::memcmp(key, h->taps_key, sizeof(key)) == 0) return GSE_SUCCESS;
  GSECHK(make_ongrid_taps(h->ct_dbl, h->X, h->Y, h->Z, h->sigma2, h->rcut2(), h->bx, h->by, h->bz,
                          h->taps, h->tapsR));
  std::memcpy(h->taps_key, key, sizeof(key));
  return GSE_SUCCESS;
}

// sym[a]: the taps of axis a are symmetric (the on-grid Gaussians always are)
const bool kSymTaps[3] = {true, true, true};
gse_status_t conv3(bool dbl, int X, int Y, int Z, const int (&R)[3], DevBuf (&taps)[3],
                   const void* in, void* out, DevBuf& tA, DevBuf& tB, hipStream_t s,
                   const bool (&sym)[3] = kSymTaps) {
  const size_t bytes = tsize(dbl) * (size_t)X * Y * Z;
  GSECHK(tA.ensure(bytes));
  GSECHK(tB.ensure(bytes));
  HIPCHK(launch_conv_pass(dbl, 0, X, Y, Z, R[0], taps[0].p, in, tA.p, s, sym[0]));
  HIPCHK(launch_conv_pass(dbl, 1, X, Y, Z, R[1], taps[1].p, tA.p, tB.p, s, sym[1]));
  HIPCHK(launch_conv_pass(dbl, 2, X, Y, Z, R[2], taps[2].p, tB.p, out, s, sym[2]));
  return GSE_SUCCESS;
}

}  // namespace

extern "C" {

const char* gse_status_string(gse_status_t s) {
  switch (s) {
    case GSE_SUCCESS: return "success";
    case GSE_ERROR_INVALID_VALUE: return "invalid value";
    case GSE_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case GSE_ERROR_HIP: return "HIP error";
    case GSE_ERROR_FFT: return "rocFFT error";
    case GSE_ERROR_BAD_PARTICLES: return "non-finite or out-of-range particles";
    case GSE_ERROR_NOT_SUPPORTED: return "not supported";
  }
  return "unknown status";
}

const char* gse_last_error(void) { return g_err.c_str(); }

int gse_version(void) { return 10000; }

gse_status_t gse_set_tiling(int bin_edge, int lds_budget_bytes) {
  if (bin_edge < 0 || lds_budget_bytes < 0 || bin_edge > 64 || lds_budget_bytes > 144 * 1024)
    return fail(GSE_ERROR_INVALID_VALUE, "bin_edge in [0,64], lds budget in [0,147456]");
  g_tuning.bin_edge = bin_edge ? bin_edge : 8;
  g_tuning.brick_budget = lds_budget_bytes ? lds_budget_bytes : 64 * 1024;
  return GSE_SUCCESS;
}

gse_status_t gse_set_tuning_param(const char* name, int value) {
  if (!name) return fail(GSE_ERROR_INVALID_VALUE, "name is NULL");
  const std::string n(name);
  Tuning& t = g_tuning;
  if (n == "bin_edge" && value >= 1 && value <= 64) t.bin_edge = value;
  else if (n == "brick_budget" && value >= 4096 && value <= 144 * 1024) t.brick_budget = value;
  else if (n == "spread_block" && (value == 0 || value == 256 || value == 512 || value == 1024))
    t.spread_block = value;
  else if (n == "gather_block" && (value == 0 || value == 256 || value == 512)) t.gather_block = value;
  else if (n == "spread_stage" && value >= 1024 && value <= 96 * 1024) t.spread_stage = value;
  else if (n == "gather_stage" && (value == 0 || (value >= 1024 && value <= 96 * 1024)))
    t.gather_stage = value;
  else if (n == "spread_lds32" && (value == 0 || value == 1)) t.spread_lds32 = value;
  else if (n == "spread_cols" && (value == 0 || value == 1)) t.spread_cols = value;
  else if (n == "spread_flush" && (value == 0 || value == 1)) t.spread_flush = value;
  else if (n == "reduce_tile" && (value == 0 || value == 1)) t.reduce_tile = value;
  else if (n == "reduce_nt" && (value == 0 || value == 1)) t.reduce_nt = value;
  else if (n == "reduce_tile2" && (value == 0 || value == 1)) t.reduce_tile2 = value;
  else if (n == "bin_agg" && (value == 0 || value == 1 || value == 2 || value == 4 || value == 8))
    t.bin_agg = value;
  else if (n == "fused_kz" && (value == 0 || value == 1)) t.fused_kz = value;
  else if (n == "zpass_cols" && (value == 4 || value == 8 || value == 16)) t.zpass_cols = value;
  else if (n == "gather_group" && (value == 16 || value == 32)) t.gather_group = value;
  else if (n == "zpass4" && (value == 0 || value == 1)) t.zpass4 = value;
  else if (n == "xy_plane" && (value == 0 || value == 1)) t.xy_plane = value;
  else if (n == "ypass4" && (value == 0 || value == 1)) t.ypass4 = value;
  else if (n == "bin_coarse_block" && (value == 256 || value == 512 || value == 1024))
    t.bin_coarse_block = value;
  else if (n == "bin_order" && (value == 0 || value == 1)) t.bin_order = value;
  else if (n == "zpass4_cols" && (value == 0 || value == 16 || value == 32))
    t.zpass4_cols = value;
  else if (n == "spread_mma" && (value == 0 || value == 1)) t.spread_mma = value;
  else return fail(GSE_ERROR_INVALID_VALUE, "unknown tuning parameter or value out of range: %s=%d",
                   name, value);
  return GSE_SUCCESS;
}

gse_status_t gse_get_tuning_param(const char* name, int* value) {
  if (!name || !value) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  const std::string n(name);
  const Tuning& t = g_tuning;
  if (n == "bin_edge") *value = t.bin_edge;
  else if (n == "brick_budget") *value = t.brick_budget;
  else if (n == "spread_block") *value = t.spread_block;
  else if (n == "gather_block") *value = t.gather_block;
  else if (n == "spread_stage") *value = t.spread_stage;
  else if (n == "gather_stage") *value = t.gather_stage;
  else if (n == "spread_lds32") *value = t.spread_lds32;
  else if (n == "spread_cols") *value = t.spread_cols;
  else if (n == "spread_flush") *value = t.spread_flush;
  else if (n == "reduce_tile") *value = t.reduce_tile;
  else if (n == "reduce_nt") *value = t.reduce_nt;
  else if (n == "reduce_tile2") *value = t.reduce_tile2;
  else if (n == "bin_agg") *value = t.bin_agg;
  else if (n == "fused_kz") *value = t.fused_kz;
  else if (n == "zpass_cols") *value = t.zpass_cols;
  else if (n == "gather_group") *value = t.gather_group;
  else if (n == "zpass4") *value = t.zpass4;
  else if (n == "xy_plane") *value = t.xy_plane;
  else if (n == "ypass4") *value = t.ypass4;
  else if (n == "zpass4_cols") *value = t.zpass4_cols;
  else if (n == "bin_coarse_block") *value = t.bin_coarse_block;
  else if (n == "bin_order") *value = t.bin_order;
  else if (n == "spread_mma") *value = t.spread_mma;
  else return fail(GSE_ERROR_INVALID_VALUE, "unknown tuning parameter: %s", name);
  return GSE_SUCCESS;
}

// ------------------------------- device memory -----------------------------
gse_status_t gse_device_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(GSE_ERROR_INVALID_VALUE, "ptr is NULL");
  *ptr = nullptr;
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return GSE_SUCCESS;
}

gse_status_t gse_device_free(void* ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return GSE_SUCCESS;
}

gse_status_t gse_memcpy_to_device(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GSE_SUCCESS;
  if (!dst || !src) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GSE_SUCCESS;
}

gse_status_t gse_memcpy_to_host(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GSE_SUCCESS;
  if (!dst || !src) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GSE_SUCCESS;
}

gse_status_t gse_memcpy_device_to_device(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GSE_SUCCESS;
  if (!dst || !src) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GSE_SUCCESS;
}

gse_status_t gse_device_clear(void* ptr, size_t bytes) {
  if (bytes == 0) return GSE_SUCCESS;
  if (!ptr) return fail(GSE_ERROR_INVALID_VALUE, "ptr is NULL");
  HIPCHK(hipMemsetAsync(ptr, 0, bytes, nullptr));
  HIPCHK(hipStreamSynchronize(nullptr));
  return GSE_SUCCESS;
}

gse_status_t gse_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return GSE_SUCCESS;
}

// ------------------------------- CpuGaussCharge ----------------------------
gse_status_t gse_gauss_charge_create(gse_gauss_charge_t* out, gse_dtype_t at, gse_dtype_t ct,
                                     int X, int Y, int Z) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (!valid_pair(at, ct)) return fail(GSE_ERROR_INVALID_VALUE, "unsupported (AT, CT) pair");
  GSECHK(check_dims(X, Y, Z));
  auto* h = new gse_gauss_charge_s();
  h->eng.at_dbl = at == GSE_FLOAT64;
  h->eng.ct_dbl = ct == GSE_FLOAT64;
  h->eng.X = X;
  h->eng.Y = Y;
  h->eng.Z = Z;
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_destroy(gse_gauss_charge_t h) {
  if (!h) return GSE_SUCCESS;
  h->eng.release();
  for (DevBuf* b : {&h->tmpA, &h->tmpB, &h->taps, &h->part, &h->dev3}) b->release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_set_stream(gse_gauss_charge_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_calc_support_size(gse_gauss_charge_t h, double rcut, double bx,
                                                double by, double bz, int* nsup) {
  if (!h || !nsup) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(check_box(bx, by, bz));
  // CpuGaussCharge.cpp:28-56, evaluated in CT like the reference
  const Engine& e = h->eng;
  const double hxd = bx / e.X, hyd = by / e.Y, hzd = bz / e.Z;
  const int nx = (int)std::ceil(rcut / hxd), ny = (int)std::ceil(rcut / hyd),
            nz = (int)std::ceil(rcut / hzd);
  int n = 0;
  if (e.ct_dbl) {
    const double rc2 = rcut * rcut;
    for (int tz = -nz; tz <= nz; tz++)
      for (int ty = -ny; ty <= ny; ty++)
        for (int tx = -nx; tx <= nx; tx++) {
          double dx = (double)(std::abs(tx) - 1) * hxd, dy = (double)(std::abs(ty) - 1) * hyd,
                 dz = (double)(std::abs(tz) - 1) * hzd;
          if (dx * dx + dy * dy + dz * dz < rc2) n++;
        }
  } else {
    const float rc2 = (float)(rcut * rcut), hx = (float)hxd, hy = (float)hyd, hz = (float)hzd;
    for (int tz = -nz; tz <= nz; tz++)
      for (int ty = -ny; ty <= ny; ty++)
        for (int tx = -nx; tx <= nx; tx++) {
          float dx = (float)(std::abs(tx) - 1) * hx, dy = (float)(std::abs(ty) - 1) * hy,
                dz = (float)(std::abs(tz) - 1) * hz;
          if (dx * dx + dy * dy + dz * dz < rc2) n++;
        }
  }
  *nsup = n;
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_spread_charge_to_grid(gse_gauss_charge_t h, double sigma,
                                                    double rcut, int n, const void* xyzq,
                                                    double bx, double by, double bz, void* rho,
                                                    int normalize) {
  if (!h || !rho) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(h->eng.bin_particles(n, xyzq, bx, by, bz, sigma, rcut, normalize != 0, h->stream));
  GSECHK(h->eng.spread_binned(sigma, normalize != 0, h->stream));
  return h->eng.convert(rho, h->stream);
}

// Allocation-free variant of gse_gauss_charge_spread_charge_to_grid: the
// same binning, spread and conversion kernels, with every buffer beyond the
// caller's particles and grid carved from the device's `scratch` array
// (noalloc.hip).  No hipMalloc/hipFree and no host allocation; the scratch
// is cleared here where the kernels need zeros.
gse_status_t gse_gauss_charge_spread_charge_to_grid_noalloc(gse_gauss_charge_t h, double sigma,
                                                            double rcut, int n,
                                                            const void* xyzq, double bx,
                                                            double by, double bz, void* rho,
                                                            int normalize) {
  if (!h || !rho) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (n < 0) return fail(GSE_ERROR_INVALID_VALUE, "numCoord must be >= 0");
  if (n > 0 && !xyzq) return fail(GSE_ERROR_INVALID_VALUE, "xyzq is NULL");
  GSECHK(check_box(bx, by, bz));
  GSECHK(check_sigma_rcut(sigma, rcut));
  const Engine& e = h->eng;
  const bool dbl = e.ct_dbl;
  const bool acc32 = !dbl && g_tuning.spread_lds32;
  const Geom g = make_geom(e.X, e.Y, e.Z, bx, by, bz, rcut);
  const BinShape b = choose_bins(g, 8, (size_t)g_tuning.brick_budget, g_tuning.bin_edge);
  const int nb = b.nbins;
  int cap_bin;
  if (b.direct) {
    cap_bin = n > 0 ? n : 1;
  } else {
    const long long c = 2LL * n / nb + 32;
    cap_bin = (int)((c + 15) / 16 * 16);
  }
  const bool flush = g_tuning.spread_flush && !g_tuning.spread_cols && brick_reduce_ok(g, b);
  const size_t slots = (size_t)nb * cap_bin, npts = (size_t)e.X * e.Y * e.Z;
  // scratch layout, each region 256-byte aligned
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t o_cnt = take(sizeof(int) * (size_t)nb), o_cnt2 = take(sizeof(int) * (size_t)nb),
               o_st = take(sizeof(BinStats)), o_st2 = take(sizeof(BinStats)),
               o_ovf = take(sizeof(int) * (size_t)(n > 0 ? n : 1)),
               o_perm = take(sizeof(int) * slots), o_loc = take(sizeof(int) * slots),
               o_sorted = take(4 * tsize(dbl) * slots), o_g64 = take(8 * npts),
               o_bricks = flush ? take((acc32 ? 4 : 8) * (size_t)b.Lx * b.Ly * b.Lz * nb) : 0,
               o_bshift = flush ? take(sizeof(int) * (size_t)nb) : 0;
  if (off > kNoAllocScratchBytes)
    return fail(GSE_ERROR_INVALID_VALUE,
                "the allocation-free spread needs %zu bytes of scratch for %d particles on a "
                "%dx%dx%d grid; its limit is %zu bytes (use "
                "gse_gauss_charge_spread_charge_to_grid)",
                off, n, e.X, e.Y, e.Z, kNoAllocScratchBytes);
  void* base = nullptr;
  HIPCHK(noalloc_scratch(&base));
  char* sp = (char*)base;
  hipStream_t s = h->stream;
  Binning w;
  w.n = n;
  w.cap_bin = cap_bin;
  w.xyzq = xyzq;
  w.counts = (int*)(sp + o_cnt);
  w.counts_other = (int*)(sp + o_cnt2);
  w.stats = (BinStats*)(sp + o_st);
  w.stats_other = (BinStats*)(sp + o_st2);
  w.ovf = (int*)(sp + o_ovf);
  w.perm = (int*)(sp + o_perm);
  w.loc = (int*)(sp + o_loc);
  w.sorted = sp + o_sorted;
  w.images = ((2 * g.nx + 1 + e.X - 1) / e.X) * ((2 * g.ny + 1 + e.Y - 1) / e.Y) *
             ((2 * g.nz + 1 + e.Z - 1) / e.Z);
  w.pref_bound = std::pow(2.0 * M_PI * sigma * sigma, -1.5) *
                 spread_weight_bound(g, sigma, normalize != 0);
  w.bricks = flush ? sp + o_bricks : nullptr;
  w.bshift = flush ? (int*)(sp + o_bshift) : nullptr;
  unsigned long long* g64 = (unsigned long long*)(sp + o_g64);
  // zeros the kernels rely on: the int64 grid (overflow / atomic sums), the
  // bin counts and statistics (fill kernels, not memsets: safe to capture)
  HIPCHK(launch_fill_u64(g64, npts, s));
  HIPCHK(launch_bin_clear(w.counts, nb, w.stats, s));
  HIPCHK(launch_bin(dbl, g, b, w, s));
  HIPCHK(launch_spread(dbl, g, b, sigma, normalize != 0, w, g64, s));
  if (flush)
    HIPCHK(launch_brick_reduce(dbl, acc32, g, b, w, g64, rho, s));
  else
    HIPCHK(launch_fixed_to_grid(dbl, g64, rho, npts, w, true, s));
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_spread_charge_on_grid(gse_gauss_charge_t h, double sigma,
                                                    double rcut, double bx, double by, double bz,
                                                    const void* rho, void* rhoS) {
  if (!h || !rho || !rhoS) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(check_box(bx, by, bz));
  GSECHK(check_sigma_rcut(sigma, rcut));
  Engine& e = h->eng;
  DevBuf taps3[3];
  int R[3];
  gse_status_t st = make_ongrid_taps(e.ct_dbl, e.X, e.Y, e.Z, sigma, rcut, bx, by, bz, taps3, R);
  if (st == GSE_SUCCESS)
    st = conv3(e.ct_dbl, e.X, e.Y, e.Z, R, taps3, rho, rhoS, h->tmpA, h->tmpB, h->stream);
  // taps are freed after the stream has consumed them
  (void)hipStreamSynchronize(h->stream);
  for (auto& t : taps3) t.release();
  return st;
}

gse_status_t gse_gauss_charge_interpolate_force(gse_gauss_charge_t h, double sigma, double rcut,
                                                int n, const void* xyzq, double bx, double by,
                                                double bz, const void* phi, void* fx, void* fy,
                                                void* fz) {
  if (!h || !phi || (n > 0 && (!fx || !fy || !fz)))
    return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(h->eng.bin_particles(n, xyzq, bx, by, bz, sigma, rcut, false, h->stream));
  GatherOut o;
  o.fx = fx;
  o.fy = fy;
  o.fz = fz;
  return h->eng.gather_binned(sigma, rcut, GF_SPHERE, phi, o, nullptr, h->stream);
}

gse_status_t gse_gauss_charge_gather(gse_gauss_charge_t h, double sigma, double rcut, int n,
                                     const void* xyzq, double bx, double by, double bz,
                                     const void* phi, const gse_gather_desc_t* d) {
  if (!h || !phi || !d) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (d->what & ~15 || d->what == 0) return fail(GSE_ERROR_INVALID_VALUE, "bad gather mask");
  GSECHK(h->eng.bin_particles(n, xyzq, bx, by, bz, sigma, rcut, false, h->stream));
  GatherOut o;
  o.fx = d->forceX;
  o.fy = d->forceY;
  o.fz = d->forceZ;
  o.cx = d->cubeX;
  o.cy = d->cubeY;
  o.cz = d->cubeZ;
  o.pot = d->potential;
  o.fscale = o.pscale = o.vscale = d->scale;
  if (d->virial) HIPCHK(launch_fill(true, d->virial, 9, h->stream));
  return h->eng.gather_binned(sigma, rcut, d->what, phi, o, d->virial, h->stream);
}

gse_status_t gse_gauss_charge_calc_electric_field_on_grid(gse_gauss_charge_t h, double bx,
                                                          double by, double bz, const void* phi,
                                                          void* ex, void* ey, void* ez) {
  if (!h || !phi || !ex || !ey || !ez) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(check_box(bx, by, bz));
  const Engine& e = h->eng;
  HIPCHK(launch_efield(e.ct_dbl, e.X, e.Y, e.Z, bx, by, bz, phi, ex, ey, ez, h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_calc_dipole_sum(gse_gauss_charge_t h, double bx, double by,
                                              double bz, const void* rho, double* dip) {
  if (!h || !rho || !dip) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(check_box(bx, by, bz));
  const Engine& e = h->eng;
  GSECHK(h->part.ensure(sizeof(double) * kRedBlocks * 3));
  GSECHK(h->dev3.ensure(sizeof(double) * 3));
  HIPCHK(launch_dipole(e.ct_dbl, e.X, e.Y, e.Z, bx, by, bz, rho, h->part.as<double>(), h->stream));
  HIPCHK(launch_reduce_partials(h->part.as<double>(), kRedBlocks, 3, 3, 1.0,
                                h->dev3.as<double>(), h->stream));
  HIPCHK(hipMemcpyAsync(dip, h->dev3.p, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_gauss_charge_invalid_count(gse_gauss_charge_t h, int* count) {
  if (!h || !count) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  return h->eng.invalid_count(h->stream, count);
}

gse_status_t gse_gauss_charge_get_layout(gse_gauss_charge_t h, int* info) {
  if (!h || !info) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  h->eng.layout(info);
  return GSE_SUCCESS;
}

// ------------------------------- CpuConvolution ----------------------------
gse_status_t gse_convolution_create(gse_convolution_t* out, gse_dtype_t t, int X, int Y, int Z) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (t != GSE_FLOAT32 && t != GSE_FLOAT64) return fail(GSE_ERROR_INVALID_VALUE, "bad dtype");
  GSECHK(check_dims(X, Y, Z));
  auto* h = new gse_convolution_s();
  h->dbl = t == GSE_FLOAT64;
  h->X = X;
  h->Y = Y;
  h->Z = Z;
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_convolution_destroy(gse_convolution_t h) {
  if (!h) return GSE_SUCCESS;
  for (auto& b : h->taps) b.release();
  h->tmpA.release();
  h->tmpB.release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_convolution_set_stream(gse_convolution_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_convolution_setup_kernel(gse_convolution_t h, int Rx, int Ry, int Rz,
                                          const void* kx, const void* ky, const void* kz) {
  if (!h || !kx || !ky || !kz) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  const int R[3] = {Rx, Ry, Rz};
  const void* k[3] = {kx, ky, kz};
  for (int a = 0; a < 3; a++) {
    if (R[a] < 0) return fail(GSE_ERROR_INVALID_VALUE, "kernel radius must be >= 0");
    const size_t bytes = tsize(h->dbl) * (2 * (size_t)R[a] + 1);
    GSECHK(h->taps[a].ensure(bytes));
    HIPCHK(hipMemcpy(h->taps[a].p, k[a], bytes, hipMemcpyHostToDevice));
    h->R[a] = R[a];
    bool sym = true;
    for (int i = 1; i <= R[a]; i++)
      sym = sym && (h->dbl ? ((const double*)k[a])[R[a] - i] == ((const double*)k[a])[R[a] + i]
                           : ((const float*)k[a])[R[a] - i] == ((const float*)k[a])[R[a] + i]);
    h->sym[a] = sym;
  }
  return GSE_SUCCESS;
}

gse_status_t gse_convolution_run(gse_convolution_t h, const void* in, void* out) {
  if (!h || !in || !out) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (h->R[0] < 0) return fail(GSE_ERROR_INVALID_VALUE, "setupKernel has not been called");
  HIPCHK(launch_conv_pass(h->dbl, 0, h->X, h->Y, h->Z, h->R[0], h->taps[0].p, in, out, h->stream,
                          h->sym[0]));
  return GSE_SUCCESS;
}

gse_status_t gse_convolution_run_xyz(gse_convolution_t h, const void* in, void* out) {
  if (!h || !in || !out) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (h->R[0] < 0) return fail(GSE_ERROR_INVALID_VALUE, "setupKernel has not been called");
  return conv3(h->dbl, h->X, h->Y, h->Z, h->R, h->taps, in, out, h->tmpA, h->tmpB, h->stream,
               h->sym);
}

// ------------------------------- CpuFFTSolver ------------------------------
gse_status_t gse_fft_solver_create(gse_fft_solver_t* out, gse_dtype_t t, gse_greens_kind_t kind,
                                   int X, int Y, int Z) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (t != GSE_FLOAT32 && t != GSE_FLOAT64) return fail(GSE_ERROR_INVALID_VALUE, "bad dtype");
  if (kind != GSE_GREENS_GSE && kind != GSE_GREENS_G2 && kind != GSE_GREENS_HERMITIAN)
    return fail(GSE_ERROR_INVALID_VALUE, "bad Green's function kind");
  GSECHK(check_dims(X, Y, Z));
  auto* h = new gse_fft_solver_s();
  h->kind = kind;
  gse_status_t st = h->ks.init(t == GSE_FLOAT64, X, Y, Z);
  if (st != GSE_SUCCESS) {
    h->ks.release();
    delete h;
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_fft_solver_destroy(gse_fft_solver_t h) {
  if (!h) return GSE_SUCCESS;
  h->ks.release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_fft_solver_set_stream(gse_fft_solver_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_fft_solver_run(gse_fft_solver_t h, void* phi, const void* rho, double sigma,
                                double bx, double by, double bz) {
  if (!h || !phi || !rho) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(check_box(bx, by, bz));
  if (!(sigma >= 0) || !std::isfinite(sigma))
    return fail(GSE_ERROR_INVALID_VALUE, "sigma must be >= 0");
  return h->ks.solve(h->kind, sigma, bx, by, bz, rho, phi, nullptr, nullptr, 1.0, h->stream);
}

gse_status_t gse_fft_solver_r2c(gse_fft_solver_t h, const void* rho, void* rhoK) {
  if (!h || !rho || !rhoK) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  return h->ks.fft.exec(true, rho, rhoK, h->stream);
}

gse_status_t gse_fft_solver_c2r(gse_fft_solver_t h, void* rhoK, void* phi) {
  if (!h || !rhoK || !phi) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  return h->ks.fft.exec(false, rhoK, phi, h->stream);
}

gse_status_t gse_fft_solver_get_form(gse_fft_solver_t h, int* form) {
  if (!h || !form) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  *form = h->ks.form();
  return GSE_SUCCESS;
}

gse_status_t gse_greens_func_mul(gse_dtype_t t, gse_greens_kind_t kind, int X, int Y, int Z,
                                 double bx, double by, double bz, double sigma, void* rhoK,
                                 void* stream) {
  if (!rhoK) return fail(GSE_ERROR_INVALID_VALUE, "rhoK is NULL");
  if (t != GSE_FLOAT32 && t != GSE_FLOAT64) return fail(GSE_ERROR_INVALID_VALUE, "bad dtype");
  if (kind != GSE_GREENS_GSE && kind != GSE_GREENS_G2 && kind != GSE_GREENS_HERMITIAN)
    return fail(GSE_ERROR_INVALID_VALUE, "bad Green's function kind");
  GSECHK(check_dims(X, Y, Z));
  GSECHK(check_box(bx, by, bz));
  const bool dbl = t == GSE_FLOAT64;
  hipStream_t s = (hipStream_t)stream;
  const size_t nk = (size_t)(X / 2 + 1) * Y * Z;
  DevBuf table, part;
  GSECHK(table.ensure(tsize(dbl) * nk));
  GSECHK(part.ensure(sizeof(double) * kKpassBlocks * 7));
  hipError_t e = launch_greens_table(dbl, kind, X, Y, Z, bx, by, bz, sigma, table.p, s);
  if (e == hipSuccess)
    e = launch_kpass(dbl, kind, X, Y, Z, bx, by, bz, sigma, rhoK, table.p, false,
                     part.as<double>(), s);
  hipError_t e2 = hipStreamSynchronize(s);
  table.release();
  part.release();
  if (e != hipSuccess) return fail(GSE_ERROR_HIP, "k-pass launch failed: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(GSE_ERROR_HIP, "k-pass failed: %s", hipGetErrorString(e2));
  return GSE_SUCCESS;
}

// ------------------------------- CpuGSEk -----------------------------------
gse_status_t gse_gsek_create(gse_gsek_t* out, gse_dtype_t at, gse_dtype_t ct, int X, int Y,
                             int Z, double sigma, double kappa, double ls, double ls1, double bx,
                             double by, double bz, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (!valid_pair(at, ct)) return fail(GSE_ERROR_INVALID_VALUE, "unsupported (AT, CT) pair");
  GSECHK(check_dims(X, Y, Z));
  GSECHK(check_box(bx, by, bz));
  if (!(kappa >= 0.0 && kappa <= 0.5))  // CpuGSE.cpp:35-36
    return fail(GSE_ERROR_INVALID_VALUE, "kappa must be in [0, 0.5]");
  if (!(sigma > 0) || !(ls1 > 0)) return fail(GSE_ERROR_INVALID_VALUE, "sigma, lambdaSigma1 > 0");
  auto* h = new gse_gsek_s();
  h->at_dbl = at == GSE_FLOAT64;
  h->ct_dbl = ct == GSE_FLOAT64;
  h->X = X;
  h->Y = Y;
  h->Z = Z;
  h->sigma = sigma;
  h->kappa = kappa;
  h->ls = ls;
  h->ls1 = ls1;
  h->ccelec = ccelec;
  h->bx = bx;
  h->by = by;
  h->bz = bz;
  h->sigma1 = sigma * std::sqrt(kappa);        // CpuGSE.cpp:42
  h->sigma2 = sigma * std::sqrt(1.0 - 2.0 * kappa);  // CpuGSE.cpp:43
  h->eng.at_dbl = h->at_dbl;
  h->eng.ct_dbl = h->ct_dbl;
  h->eng.X = X;
  h->eng.Y = Y;
  h->eng.Z = Z;
  const size_t gb = tsize(h->ct_dbl) * (size_t)X * Y * Z;
  gse_status_t st = h->ks.init(h->ct_dbl, X, Y, Z);
  if (st == GSE_SUCCESS) st = h->rhoS.ensure(gb);
  if (st == GSE_SUCCESS) st = h->phiM.ensure(gb);
  if (st == GSE_SUCCESS) st = h->dout.ensure(sizeof(double) * 16);
  if (st == GSE_SUCCESS) st = h->part.ensure(sizeof(double) * kRedBlocks * 8);
  if (st != GSE_SUCCESS) {
    gse_gsek_destroy(h);
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_destroy(gse_gsek_t h) {
  if (!h) return GSE_SUCCESS;
  for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
  h->eng.release();
  h->ks.release();
  for (DevBuf* b : {&h->rhoS, &h->phiM, &h->rhoR, &h->tmp, &h->tmp2, &h->dout, &h->part})
    b->release();
  for (auto& t : h->taps) t.release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_set_stream(gse_gsek_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_set_box_size(gse_gsek_t h, double bx, double by, double bz) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  GSECHK(check_box(bx, by, bz));
  h->bx = bx;
  h->by = by;
  h->bz = bz;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_get_params(gse_gsek_t h, double* s1, double* s2, double* ls1) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (s1) *s1 = h->sigma1;
  if (s2) *s2 = h->sigma2;
  if (ls1) *ls1 = h->ls1;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_spread_charge1(gse_gsek_t h, int n, const void* xyzq) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  GSECHK(h->eng.bin_particles(n, xyzq, h->bx, h->by, h->bz, h->sigma1, h->rcut1(), false,
                              h->stream));
  GSECHK(h->eng.spread_binned(h->sigma1, false, h->stream));
  return h->eng.convert(h->rhoS.p, h->stream);
}

gse_status_t gse_gsek_spread_charge2(gse_gsek_t h, void* rhoOut) {
  if (!h || !rhoOut) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(gsek_taps(h));
  return conv3(h->ct_dbl, h->X, h->Y, h->Z, h->tapsR, h->taps, h->rhoS.p, rhoOut, h->tmp, h->tmp2,
               h->stream);
}

gse_status_t gse_gsek_solve_poisson(gse_gsek_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return h->ks.solve(GSE_GREENS_GSE, h->sigma2, h->bx, h->by, h->bz, h->rhoS.p, h->phiM.p,
                     nullptr, nullptr, h->ccelec, h->stream);
}

gse_status_t gse_gsek_calculate_energy(gse_gsek_t h, double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  const double vcell = (h->bx / h->X) * (h->by / h->Y) * (h->bz / h->Z);
  HIPCHK(launch_dot(h->ct_dbl, h->rhoS.p, h->phiM.p, (size_t)h->X * h->Y * h->Z,
                    h->part.as<double>(), h->stream));
  HIPCHK(launch_reduce_partials(h->part.as<double>(), kRedBlocks, 1, 1, 0.5 * h->ccelec * vcell,
                                h->dout.as<double>(), h->stream));
  HIPCHK(hipMemcpyAsync(energy, h->dout.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_interpolate_force(gse_gsek_t h, int n, const void* xyzq, void* fx, void* fy,
                                        void* fz) {
  if (!h || (n > 0 && (!fx || !fy || !fz))) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(h->eng.bin_particles(n, xyzq, h->bx, h->by, h->bz, h->sigma1, h->rcut1(), false,
                              h->stream));
  GatherOut o;
  o.fx = fx;
  o.fy = fy;
  o.fz = fz;
  o.fscale = h->ccelec;
  return h->eng.gather_binned(h->sigma1, h->rcut1(), GF_SPHERE, h->phiM.p, o, nullptr, h->stream);
}

gse_status_t gse_gsek_get_layout(gse_gsek_t h, int* info) {
  if (!h || !info) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  h->eng.layout(info, h->ks.form());
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_get_rho_s(gse_gsek_t h, void** p) {
  if (!h || !p) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  *p = h->rhoS.p;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_get_phi(gse_gsek_t h, void** p) {
  if (!h || !p) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  *p = h->phiM.p;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_copy_grid(gse_gsek_t h, int which, void* dst) {
  if (!h || !dst || (which != 0 && which != 1))
    return fail(GSE_ERROR_INVALID_VALUE, "bad argument");
  const size_t bytes = tsize(h->ct_dbl) * (size_t)h->X * h->Y * h->Z;
  HIPCHK(hipMemcpyAsync(dst, which == 0 ? h->rhoS.p : h->phiM.p, bytes, hipMemcpyDeviceToDevice,
                        h->stream));
  return GSE_SUCCESS;
}

static gse_status_t mark(gse_gsek_t h);
static gse_status_t stage_mark(gse_gsek_s* h) { return mark(h); }

static gse_status_t mark(gse_gsek_t h) {
  if (!h->profiling) return GSE_SUCCESS;
  if (h->ev_used == h->events.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    h->events.push_back(e);
  }
  HIPCHK(hipEventRecord(h->events[h->ev_used++], h->stream));
  return GSE_SUCCESS;
}

static gse_status_t gsek_run(gse_gsek_t h, int n, const void* xyzq, int mode, int flags,
                             GatherOut& o, double* energy, double* virial) {
  hipStream_t s = h->stream;
  const double rc1 = h->rcut1();
  GSECHK(mark(h));
  GSECHK(h->eng.bin_particles(n, xyzq, h->bx, h->by, h->bz, h->sigma1, rc1, false, s));
  GSECHK(mark(h));
  GSECHK(h->eng.spread_binned(h->sigma1, false, s));
  GSECHK(mark(h));
  GSECHK(h->eng.convert(h->rhoS.p, s));
  GSECHK(mark(h));
  double* eout = energy ? energy : h->dout.as<double>();
  if (mode == 0) {
    GSECHK(h->ks.solve(GSE_GREENS_GSE, h->sigma2, h->bx, h->by, h->bz, h->rhoS.p, h->phiM.p, eout,
                       virial, h->ccelec, s, h->profiling ? h : nullptr, flags ? &o : nullptr));
  } else {
    GSECHK(gsek_taps(h));
    GSECHK(h->rhoR.ensure(tsize(h->ct_dbl) * (size_t)h->X * h->Y * h->Z));
    GSECHK(conv3(h->ct_dbl, h->X, h->Y, h->Z, h->tapsR, h->taps, h->rhoS.p, h->rhoR.p, h->tmp,
                 h->tmp2, s));
    GSECHK(h->ks.solve(GSE_GREENS_GSE, 0.0, h->bx, h->by, h->bz, h->rhoR.p, h->phiM.p, nullptr,
                       nullptr, h->ccelec, s, h->profiling ? h : nullptr));
    const double vcell = (h->bx / h->X) * (h->by / h->Y) * (h->bz / h->Z);
    HIPCHK(launch_dot(h->ct_dbl, h->rhoS.p, h->phiM.p, (size_t)h->X * h->Y * h->Z,
                      h->part.as<double>(), s));
    HIPCHK(launch_reduce_partials(h->part.as<double>(), kRedBlocks, 1, 1,
                                  0.5 * h->ccelec * vcell, eout, s));
  }
  if (flags) GSECHK(h->eng.gather_binned(h->sigma1, rc1, flags, h->phiM.p, o,
                                         (virial && mode == 0) ? virial : nullptr, s));
  GSECHK(mark(h));
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_set_profiling(gse_gsek_t h, int enable) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->profiling = enable != 0;
  h->ev_used = 0;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_stage_times(gse_gsek_t h, double* ms, int* steps) {
  if (!h || !ms || !steps) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  const int per = GSE_NUM_STAGES + 1;
  const int ns = (int)(h->ev_used / per);
  for (int k = 0; k < GSE_NUM_STAGES; k++) ms[k] = 0.0;
  if (ns > 0) HIPCHK(hipEventSynchronize(h->events[(size_t)ns * per - 1]));
  for (int st = 0; st < ns; st++)
    for (int k = 0; k < GSE_NUM_STAGES; k++) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, h->events[(size_t)st * per + k],
                                 h->events[(size_t)st * per + k + 1]));
      ms[k] += t;
    }
  if (ns > 0)
    for (int k = 0; k < GSE_NUM_STAGES; k++) ms[k] /= ns;
  *steps = ns;
  h->ev_used = 0;
  return GSE_SUCCESS;
}

gse_status_t gse_gsek_step(gse_gsek_t h, int n, const void* xyzq, const gse_step_desc_t* d) {
  if (!h || !d) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (d->mode != 0 && d->mode != 1) return fail(GSE_ERROR_INVALID_VALUE, "mode must be 0 or 1");
  if (d->mode == 1 && d->virial)
    return fail(GSE_ERROR_NOT_SUPPORTED, "the virial is defined for mode K only");
  int flags = 0;
  if (d->forceX || d->forceY || d->forceZ) flags |= GF_SPHERE;
  if (d->cubeX || d->cubeY || d->cubeZ) flags |= GF_CUBE;
  if (d->potential) flags |= GF_POT;
  if (d->virial) flags |= GF_VIRIAL;
  if ((flags & GF_SPHERE) && !(d->forceX && d->forceY && d->forceZ))
    return fail(GSE_ERROR_INVALID_VALUE, "forceX/Y/Z must be all set or all NULL");
  if ((flags & GF_CUBE) && !(d->cubeX && d->cubeY && d->cubeZ))
    return fail(GSE_ERROR_INVALID_VALUE, "cubeX/Y/Z must be all set or all NULL");
  GatherOut o;
  o.fx = d->forceX;
  o.fy = d->forceY;
  o.fz = d->forceZ;
  o.cx = d->cubeX;
  o.cy = d->cubeY;
  o.cz = d->cubeZ;
  o.pot = d->potential;
  o.fscale = o.pscale = o.vscale = h->ccelec;
  return gsek_run(h, n, xyzq, d->mode, flags, o, d->energy, d->virial);
}

gse_status_t gse_gsek_adjoint(gse_gsek_t h, int n, const void* xyzq, const double* grad,
                              void* dEdq, void* dEdx, void* dEdy, void* dEdz, double* virial) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if ((dEdx || dEdy || dEdz) && !(dEdx && dEdy && dEdz))
    return fail(GSE_ERROR_INVALID_VALUE, "dEdx/y/z must be all set or all NULL");
  int flags = 0;
  if (dEdx) flags |= GF_CUBE;
  if (dEdq) flags |= GF_POT;
  if (virial) flags |= GF_VIRIAL;
  GatherOut o;
  o.cx = dEdx;  // the cube force is -dE/dr: the negative scale gives dE/dr
  o.cy = dEdy;
  o.cz = dEdz;
  o.pot = dEdq;
  o.fscale = -h->ccelec;
  o.pscale = h->ccelec;
  o.vscale = h->ccelec;
  o.dscale = grad;
  return gsek_run(h, n, xyzq, 0, flags, o, nullptr, virial);
}

gse_status_t gse_gsek_corrections(gse_gsek_t h, int n, const void* xyzq, double* out) {
  if (!h || !out || (n > 0 && !xyzq)) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  HIPCHK(launch_charge_moments(h->ct_dbl, xyzq, n, h->part.as<double>(), h->stream));
  HIPCHK(launch_corrections_finalize(h->part.as<double>(), kRedBlocks, h->sigma, h->bx, h->by,
                                     h->bz, h->ccelec, out, h->stream));
  return GSE_SUCCESS;
}

// ------------------------------- SPME (CudaPMERecip) ----------------------
}  // extern "C"

namespace {

// calc_prefac / dftmod (CudaPMERecip.cu:3385-3478): 1/|sum_j M_n(j) e^{2 pi i k j/M}|^2
// per axis, with the reference's replacement of a vanishing modulus by the
// mean of its neighbours; FP64 on the host
void pme_moduli(int order, int nfft, double* out) {
  double th[16];
  th[order - 1] = 0.0;
  th[1] = 0.0;
  th[0] = 1.0;
  for (int k = 3; k <= order - 1; k++) {
    const double div = 1.0 / (double)(k - 1);
    th[k - 1] = 0.0;
    for (int j = 1; j <= k - 2; j++) th[k - j - 1] = div * (j * th[k - j - 2] + (k - j) * th[k - j - 1]);
    th[0] = div * th[0];
  }
  {
    const double div = 1.0 / (double)(order - 1);
    th[order - 1] = 0.0;
    for (int j = 1; j <= order - 2; j++)
      th[order - j - 1] = div * (j * th[order - j - 2] + (order - j) * th[order - j - 1]);
    th[0] = div * th[0];
  }
  std::vector<double> arr((size_t)nfft + order + 2, 0.0);
  for (int i = 2; i <= order + 1; i++) arr[i - 1] = th[i - 2];
  const double r = 2.0 * M_PI / (double)nfft;
  for (int k = 1; k <= nfft; k++) {
    double s1 = 0.0, s2 = 0.0;
    const double a1 = (k - 1) * r;
    for (int j = 1; j < nfft; j++) {
      const double a = a1 * (j - 1);
      s1 += arr[j - 1] * std::cos(a);
      s2 += arr[j - 1] * std::sin(a);
    }
    out[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= nfft; k++)
    if (out[k - 1] < 1.0e-10) out[k - 1] = 0.5 * (out[(k - 2 + nfft) % nfft] + out[k % nfft]);
  for (int k = 0; k < nfft; k++) out[k] = 1.0 / out[k];
}

gse_status_t pme_recip3(const double* recip, double* r3) {
  if (!recip) return fail(GSE_ERROR_INVALID_VALUE, "recip is NULL");
  if (recip[1] != 0.0 || recip[2] != 0.0 || recip[3] != 0.0 || recip[5] != 0.0 ||
      recip[6] != 0.0 || recip[7] != 0.0)
    return fail(GSE_ERROR_NOT_SUPPORTED, "only orthorhombic boxes are supported (as the reference)");
  if (!(recip[0] > 0) || !(recip[4] > 0) || !(recip[8] > 0) || !std::isfinite(recip[0]) ||
      !std::isfinite(recip[4]) || !std::isfinite(recip[8]))
    return fail(GSE_ERROR_INVALID_VALUE, "recip diagonal must be positive and finite");
  r3[0] = recip[0];
  r3[1] = recip[4];
  r3[2] = recip[8];
  return GSE_SUCCESS;
}

// the int64 -> grid conversion reads max |q| and the count from a Binning
Binning pme_binning(gse_pme_recip_s* h, int n) {
  Binning w;
  w.n = n;
  w.stats = h->stats.as<BinStats>();
  w.images = 1;
  w.pref_bound = 1.0;
  return w;
}

constexpr double kPmeCcelec = 332.0716;  // CudaPMERecip.h:25

}  // namespace

extern "C" {

gse_status_t gse_pme_recip_create(gse_pme_recip_t* out, gse_dtype_t at, gse_dtype_t ct,
                                  int nfftx, int nffty, int nfftz, int order) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (!valid_pair(at, ct)) return fail(GSE_ERROR_INVALID_VALUE, "unsupported (AT, CT) pair");
  GSECHK(check_dims(nfftx, nffty, nfftz));
  if (order != 4 && order != 6 && order != 8)
    return fail(GSE_ERROR_INVALID_VALUE, "order must be 4, 6 or 8 (CudaPMERecip.cu:2841-2893)");
  if (nfftx < order || nffty < order || nfftz < order)
    return fail(GSE_ERROR_INVALID_VALUE, "every grid dimension must be at least the order");
  auto* h = new gse_pme_recip_s();
  h->at_dbl = at == GSE_FLOAT64;
  h->ct_dbl = ct == GSE_FLOAT64;
  h->X = nfftx;
  h->Y = nffty;
  h->Z = nfftz;
  h->order = order;
  h->fft = new KSolverFFTOnly();
  const size_t np = (size_t)nfftx * nffty * nfftz;
  const size_t nk = (size_t)(nfftx / 2 + 1) * nffty * nfftz;
  gse_status_t st = h->fft->fft.init(h->ct_dbl, nfftx, nffty, nfftz);
  if (st == GSE_SUCCESS) st = h->grid64.ensure(8 * np);
  if (st == GSE_SUCCESS) st = h->grid.ensure(tsize(h->ct_dbl) * np);
  if (st == GSE_SUCCESS) st = h->spec.ensure(2 * tsize(h->ct_dbl) * nk);
  if (st == GSE_SUCCESS) st = h->prefac.ensure(sizeof(double) * (size_t)(nfftx + nffty + nfftz));
  if (st == GSE_SUCCESS) st = h->partials.ensure(sizeof(double) * 7 * kPmeSumBlocks);
  if (st == GSE_SUCCESS) st = h->stats.ensure(sizeof(BinStats));
  if (st == GSE_SUCCESS) st = h->tmp2.ensure(sizeof(double) * (kRedBlocks * 2 + 2));
  if (st == GSE_SUCCESS) {
    std::vector<double> pf((size_t)nfftx + nffty + nfftz);
    pme_moduli(order, nfftx, pf.data());
    pme_moduli(order, nffty, pf.data() + nfftx);
    pme_moduli(order, nfftz, pf.data() + nfftx + nffty);
    hipError_t e = hipMemcpy(h->prefac.p, pf.data(), sizeof(double) * pf.size(),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_fill_u64(h->grid64.as<unsigned long long>(), np, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) st = fail(GSE_ERROR_HIP, "SPME set-up failed: %s", hipGetErrorString(e));
  }
  if (st != GSE_SUCCESS) {
    gse_pme_recip_destroy(h);
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_destroy(gse_pme_recip_t h) {
  if (!h) return GSE_SUCCESS;
  if (h->fft) {
    h->fft->fft.release();
    delete h->fft;
  }
  for (DevBuf* b : {&h->grid64, &h->grid, &h->spec, &h->prefac, &h->partials, &h->stats, &h->tmp2})
    b->release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_set_stream(gse_pme_recip_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_spread_charge(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                         const double* recip) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (ncoord < 0 || (ncoord > 0 && !xyzq)) return fail(GSE_ERROR_INVALID_VALUE, "bad particles");
  double r3[3];
  GSECHK(pme_recip3(recip, r3));
  std::memcpy(h->recip3, r3, sizeof(r3));
  h->n_last = ncoord;
  hipStream_t s = h->stream;
  HIPCHK(launch_pme_qmax(h->ct_dbl, xyzq, ncoord, r3, h->stats.as<BinStats>(), s));
  HIPCHK(launch_pme_spread(h->ct_dbl, h->order, xyzq, ncoord, h->X, h->Y, h->Z, r3,
                           h->stats.as<BinStats>(), h->grid64.as<unsigned long long>(), s));
  const Binning w = pme_binning(h, ncoord);
  HIPCHK(launch_fixed_to_grid(h->ct_dbl, h->grid64.as<unsigned long long>(), h->grid.p,
                              (size_t)h->X * h->Y * h->Z, w, true, s));
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_r2c_fft(gse_pme_recip_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return h->fft->fft.exec(true, h->grid.p, h->spec.p, h->stream);
}

gse_status_t gse_pme_recip_scalar_sum(gse_pme_recip_t h, const double* recip, double kappa,
                                      int calc_energy, int calc_virial, double* energy,
                                      double* virial) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (!(kappa > 0) || !std::isfinite(kappa)) return fail(GSE_ERROR_INVALID_VALUE, "kappa > 0");
  if ((calc_energy && !energy) || (calc_virial && !virial))
    return fail(GSE_ERROR_INVALID_VALUE, "energy/virial output is NULL");
  double r3[3];
  GSECHK(pme_recip3(recip, r3));
  const bool want = calc_energy || calc_virial;
  HIPCHK(launch_pme_scalar_sum(h->ct_dbl, h->X, h->Y, h->Z, r3, kappa, h->prefac.as<double>(),
                               h->spec.p, want, h->partials.as<double>(), h->stream));
  if (want)
    HIPCHK(launch_pme_finalize(h->partials.as<double>(), kPmeCcelec, calc_energy ? energy : nullptr,
                               calc_virial ? virial : nullptr, h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_c2r_fft(gse_pme_recip_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return h->fft->fft.exec(false, h->spec.p, h->grid.p, h->stream);
}

gse_status_t gse_pme_recip_gather_force(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                        const double* recip, int stride, void* force) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (ncoord < 0 || (ncoord > 0 && (!xyzq || !force)))
    return fail(GSE_ERROR_INVALID_VALUE, "bad particles or force array");
  if (stride < ncoord) return fail(GSE_ERROR_INVALID_VALUE, "stride must be >= ncoord");
  double r3[3];
  GSECHK(pme_recip3(recip, r3));
  const size_t as = tsize(h->at_dbl);
  char* f = (char*)force;
  HIPCHK(launch_pme_gather(h->ct_dbl, h->at_dbl, h->order, xyzq, ncoord, h->X, h->Y, h->Z, r3,
                           kPmeCcelec, 1.0, h->grid.p, f, f + as * (size_t)stride,
                           f + 2 * as * (size_t)stride, h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_calc_self_energy(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                            double kappa, double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (ncoord < 0 || (ncoord > 0 && !xyzq)) return fail(GSE_ERROR_INVALID_VALUE, "bad particles");
  if (!(kappa > 0)) return fail(GSE_ERROR_INVALID_VALUE, "kappa > 0");
  double* t = h->tmp2.as<double>();
  HIPCHK(launch_charge_moments(h->ct_dbl, xyzq, ncoord, t, h->stream));
  // -kappa ccelec/sqrt(pi) sum q^2 (CudaPMERecip.cu:2359-2380) is the
  // Gaussian self term with sigma = 1/(sqrt(2) kappa)
  HIPCHK(launch_corrections_finalize(t, kRedBlocks, 1.0 / (std::sqrt(2.0) * kappa), 1.0, 1.0, 1.0,
                                     kPmeCcelec, t + 2 * kRedBlocks, h->stream));
  HIPCHK(hipMemcpyAsync(energy, t + 2 * kRedBlocks, sizeof(double), hipMemcpyDeviceToDevice,
                        h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_step(gse_pme_recip_t h, const void* xyzq, int ncoord, double boxx,
                                double boxy, double boxz, double kappa, double* energy,
                                double* virial, void* fx, void* fy, void* fz) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  GSECHK(check_box(boxx, boxy, boxz));
  if ((fx || fy || fz) && !(fx && fy && fz))
    return fail(GSE_ERROR_INVALID_VALUE, "fx/fy/fz must be all set or all NULL");
  const double recip[9] = {1.0 / boxx, 0, 0, 0, 1.0 / boxy, 0, 0, 0, 1.0 / boxz};
  GSECHK(gse_pme_recip_spread_charge(h, xyzq, ncoord, recip));
  GSECHK(gse_pme_recip_r2c_fft(h));
  GSECHK(gse_pme_recip_scalar_sum(h, recip, kappa, energy != nullptr, virial != nullptr, energy,
                                  virial));
  if (fx) {
    GSECHK(gse_pme_recip_c2r_fft(h));
    HIPCHK(launch_pme_gather(h->ct_dbl, h->at_dbl, h->order, xyzq, ncoord, h->X, h->Y, h->Z,
                             h->recip3, kPmeCcelec, -1.0, h->grid.p, fx, fy, fz, h->stream));
  }
  return GSE_SUCCESS;
}

gse_status_t gse_pme_recip_get_grid(gse_pme_recip_t h, void** grid) {
  if (!h || !grid) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  *grid = h->grid.p;
  return GSE_SUCCESS;
}

// ------------------------------- Ewald direct (CpuEwaldDirect) --------------
gse_status_t gse_ewald_direct_create(gse_ewald_direct_t* out, double sigma, double rcut,
                                     double boxx, double boxy, double boxz, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  GSECHK(check_box(boxx, boxy, boxz));
  GSECHK(check_sigma_rcut(sigma, rcut));
  auto* h = new gse_ewald_direct_s();
  h->sigma = sigma;
  h->rcut = rcut;
  h->ccelec = ccelec;
  h->box[0] = boxx;
  h->box[1] = boxy;
  h->box[2] = boxz;
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_direct_destroy(gse_ewald_direct_t h) {
  if (!h) return GSE_SUCCESS;
  for (DevBuf* b : {&h->xofs, &h->xlist, &h->cell_of, &h->counts, &h->starts, &h->fill, &h->tmp,
                    &h->sidx, &h->sx, &h->partials, &h->self2, &h->mom})
    b->release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_direct_set_stream(gse_ewald_direct_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_direct_setup_exclusions(gse_ewald_direct_t h, int n_iblo14,
                                               const int* iblo14, const int* inb14) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (n_iblo14 < 0 || (n_iblo14 > 0 && (!iblo14 || !inb14)))
    return fail(GSE_ERROR_INVALID_VALUE, "bad exclusion arrays");
  if (n_iblo14 == 0) {
    h->have_excl = false;
    h->n_excl_atoms = 0;
    return GSE_SUCCESS;
  }
  // CpuEwaldDirect.cpp:31-40: atom j excludes inb14[iblo14[j-1] .. iblo14[j]) (1-based,
  // sign ignored) among the atoms after it; both ends of each pair get it here
  const int ntot = iblo14[n_iblo14 - 1];
  std::vector<std::vector<int>> lists(n_iblo14);
  int prev = 0;
  for (int j = 0; j < n_iblo14; j++) {
    const int end = iblo14[j];
    if (end < prev || end > ntot)
      return fail(GSE_ERROR_INVALID_VALUE, "iblo14 must be non-decreasing running ends");
    for (int k = prev; k < end; k++) {
      const int i = std::abs(inb14[k]) - 1;
      if (i < 0 || i >= n_iblo14) return fail(GSE_ERROR_INVALID_VALUE, "inb14 atom out of range");
      if (i > j) {  // the reference only tests partners after j
        lists[j].push_back(i);
        lists[i].push_back(j);
      }
    }
    prev = end;
  }
  std::vector<int> ofs(n_iblo14 + 1, 0), flat;
  for (int j = 0; j < n_iblo14; j++) {
    std::sort(lists[j].begin(), lists[j].end());
    lists[j].erase(std::unique(lists[j].begin(), lists[j].end()), lists[j].end());
    flat.insert(flat.end(), lists[j].begin(), lists[j].end());
    ofs[j + 1] = (int)flat.size();
  }
  GSECHK(h->xofs.ensure(sizeof(int) * ofs.size()));
  GSECHK(h->xlist.ensure(sizeof(int) * std::max<size_t>(1, flat.size())));
  HIPCHK(hipMemcpy(h->xofs.p, ofs.data(), sizeof(int) * ofs.size(), hipMemcpyHostToDevice));
  if (!flat.empty())
    HIPCHK(hipMemcpy(h->xlist.p, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
  h->have_excl = true;
  h->n_excl_atoms = n_iblo14;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_direct_calc_force_energy(gse_ewald_direct_t h, int n, const void* xyzq,
                                                double* fx, double* fy, double* fz,
                                                double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (n < 0 || (n > 0 && (!xyzq || !fx || !fy || !fz)))
    return fail(GSE_ERROR_INVALID_VALUE, "bad particle or force arrays");
  if (h->have_excl && h->n_excl_atoms != n)  // CpuEwaldDirect.cpp:47
    return fail(GSE_ERROR_INVALID_VALUE, "the exclusion list covers %d atoms, the call %d",
                h->n_excl_atoms, n);
  EwCells c;
  c.ncell = 1;
  for (int a = 0; a < 3; a++) {
    c.box[a] = h->box[a];
    int nc = (int)std::floor(h->box[a] / h->rcut);
    if (nc > 128) nc = 128;
    c.nc[a] = nc >= 3 ? nc : 1;
    c.ncell *= c.nc[a];
  }
  EwPair p;
  p.rcut2 = h->rcut * h->rcut;
  p.ccelec = h->ccelec;
  p.inv_sqrt2sigma = 1.0 / (std::sqrt(2.0) * h->sigma);
  p.inv_2sigmasq = 1.0 / (2.0 * h->sigma * h->sigma);
  p.two_sqrtpi = 2.0 / std::sqrt(M_PI);
  const size_t n1 = (size_t)(n > 0 ? n : 1);
  GSECHK(h->cell_of.ensure(sizeof(int) * n1));
  GSECHK(h->tmp.ensure(sizeof(int) * n1));
  GSECHK(h->sidx.ensure(sizeof(int) * n1));
  GSECHK(h->sx.ensure(4 * sizeof(double) * n1));
  GSECHK(h->counts.ensure(sizeof(int) * c.ncell));
  GSECHK(h->starts.ensure(sizeof(int) * c.ncell));
  GSECHK(h->fill.ensure(sizeof(int) * c.ncell));
  GSECHK(h->partials.ensure(sizeof(double) * c.ncell));
  GSECHK(h->mom.ensure(sizeof(double) * kRedBlocks * 2));
  GSECHK(h->self2.ensure(sizeof(double) * 2));
  hipStream_t s = h->stream;
  HIPCHK(launch_ew_cells(xyzq, n, c, h->cell_of.as<int>(), h->counts.as<int>(),
                         h->starts.as<int>(), h->fill.as<int>(), h->tmp.as<int>(),
                         h->sidx.as<int>(), h->sx.p, s));
  HIPCHK(launch_ew_force(c, p, h->starts.as<int>(), h->counts.as<int>(), h->sidx.as<int>(),
                         h->sx.p, h->have_excl ? h->xofs.as<int>() : nullptr,
                         h->have_excl ? h->xlist.as<int>() : nullptr, fx, fy, fz,
                         h->partials.as<double>(), s));
  // self energy -ccelec sum q^2 / sqrt(2 pi sigma^2) (CpuEwaldDirect.cpp:124-128)
  HIPCHK(launch_charge_moments(true, xyzq, n, h->mom.as<double>(), s));
  HIPCHK(launch_corrections_finalize(h->mom.as<double>(), kRedBlocks, h->sigma, 1.0, 1.0, 1.0,
                                     h->ccelec, h->self2.as<double>(), s));
  HIPCHK(launch_ew_finalize(h->partials.as<double>(), c.ncell, h->self2.as<double>(), energy, s));
  return GSE_SUCCESS;
}

// ------------------------------- Ewald reciprocal (CpuEwaldRecip) --------
gse_status_t gse_ewald_recip_create(gse_ewald_recip_t* out, double sigma, int kcut, double boxx,
                                    double boxy, double boxz, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  GSECHK(check_box(boxx, boxy, boxz));
  if (!(sigma >= 0) || !std::isfinite(sigma)) return fail(GSE_ERROR_INVALID_VALUE, "sigma >= 0");
  if (kcut < 0 || kcut > 400) return fail(GSE_ERROR_INVALID_VALUE, "kcut must be in [0, 400]");
  auto* h = new gse_ewald_recip_s();
  h->sigma = sigma;
  h->kcut = kcut;
  h->ccelec = ccelec;
  h->box[0] = boxx;
  h->box[1] = boxy;
  h->box[2] = boxz;
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_recip_destroy(gse_ewald_recip_t h) {
  if (!h) return GSE_SUCCESS;
  h->sk.release();
  h->partials.release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_recip_set_stream(gse_ewald_recip_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_ewald_recip_calc_force_energy(gse_ewald_recip_t h, int n, const void* xyzq,
                                               double* fx, double* fy, double* fz,
                                               double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  if (n < 0 || (n > 0 && !xyzq)) return fail(GSE_ERROR_INVALID_VALUE, "bad particles");
  if ((fx || fy || fz) && !(fx && fy && fz))
    return fail(GSE_ERROR_INVALID_VALUE, "fx/fy/fz must be all set or all NULL");
  ErArgs a;
  a.kcut = h->kcut;
  const long long n1 = 2LL * h->kcut + 1;
  a.ncube = (int)(n1 * n1 * n1);
  for (int d = 0; d < 3; d++) a.ib[d] = 1.0 / h->box[d];
  a.pref = 2.0 * M_PI * a.ib[0] * a.ib[1] * a.ib[2] * h->ccelec;  // CpuEwaldRecip.cpp:95
  a.half_sigmasq = 0.5 * h->sigma * h->sigma;
  GSECHK(h->sk.ensure(3 * sizeof(double) * (size_t)a.ncube));
  GSECHK(h->partials.ensure(sizeof(double) * (size_t)ewald_recip_blocks(a)));
  HIPCHK(launch_ewald_recip(a, xyzq, n, h->sk.as<double>(), h->partials.as<double>(), fx, fy, fz,
                            energy, h->stream));
  return GSE_SUCCESS;
}

// ------------------------------- GSE-r (CpuGSEr) ---------------------------
gse_status_t gse_gser_create(gse_gser_t* out, gse_dtype_t at, gse_dtype_t ct, int X, int Y,
                             int Z, double sigma, double kappa, double ls, double ls1, double bx,
                             double by, double bz, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  gse_gsek_t core = nullptr;
  GSECHK(gse_gsek_create(&core, at, ct, X, Y, Z, sigma, kappa, ls, ls1, bx, by, bz, ccelec));
  auto* h = new gse_gser_s();
  h->core = core;
  const gse_status_t st = h->rho.ensure(tsize(core->ct_dbl) * (size_t)X * Y * Z);
  if (st != GSE_SUCCESS) {
    gse_gser_destroy(h);
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_gser_destroy(gse_gser_t h) {
  if (!h) return GSE_SUCCESS;
  gse_gsek_destroy(h->core);
  h->rho.release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_gser_set_stream(gse_gser_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return gse_gsek_set_stream(h->core, s);
}

gse_status_t gse_gser_spread_charge1(gse_gser_t h, int n, const void* xyzq) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  gse_gsek_s* c = h->core;
  GSECHK(c->eng.bin_particles(n, xyzq, c->bx, c->by, c->bz, c->sigma1, c->rcut1(), false,
                              c->stream));
  GSECHK(c->eng.spread_binned(c->sigma1, false, c->stream));
  return c->eng.convert(h->rho.p, c->stream);
}

gse_status_t gse_gser_spread_charge2(gse_gser_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  gse_gsek_s* c = h->core;
  GSECHK(gsek_taps(c));
  return conv3(c->ct_dbl, c->X, c->Y, c->Z, c->tapsR, c->taps, h->rho.p, c->rhoS.p, c->tmp,
               c->tmp2, c->stream);
}

gse_status_t gse_gser_solve_poisson(gse_gser_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  gse_gsek_s* c = h->core;
  return c->ks.solve(GSE_GREENS_HERMITIAN, 0.0, c->bx, c->by, c->bz, c->rhoS.p, c->phiM.p,
                     nullptr, nullptr, c->ccelec, c->stream);
}

// 1/2 ccelec h^3 sum a*b into dst (device)
static gse_status_t grid_energy(gse_gsek_s* c, const void* a, const void* b, double scale,
                                double* dst) {
  const double vcell = (c->bx / c->X) * (c->by / c->Y) * (c->bz / c->Z);
  HIPCHK(launch_dot(c->ct_dbl, a, b, (size_t)c->X * c->Y * c->Z, c->part.as<double>(), c->stream));
  HIPCHK(launch_reduce_partials(c->part.as<double>(), kRedBlocks, 1, 1,
                                scale * 0.5 * c->ccelec * vcell, dst, c->stream));
  return GSE_SUCCESS;
}

static gse_status_t host_energy(gse_gsek_s* c, double* energy) {
  HIPCHK(hipMemcpyAsync(energy, c->dout.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_gser_calculate_energy(gse_gser_t h, double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  GSECHK(grid_energy(h->core, h->rho.p, h->core->phiM.p, 1.0, h->core->dout.as<double>()));
  return host_energy(h->core, energy);
}

gse_status_t gse_gser_interpolate_force(gse_gser_t h, int n, const void* xyzq, void* fx,
                                        void* fy, void* fz) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return gse_gsek_interpolate_force(h->core, n, xyzq, fx, fy, fz);
}

gse_status_t gse_gser_copy_grid(gse_gser_t h, int which, void* dst) {
  if (!h || !dst || which < 0 || which > 2) return fail(GSE_ERROR_INVALID_VALUE, "bad argument");
  gse_gsek_s* c = h->core;
  const void* src = which == 0 ? h->rho.p : (which == 1 ? c->rhoS.p : c->phiM.p);
  HIPCHK(hipMemcpyAsync(dst, src, tsize(c->ct_dbl) * (size_t)c->X * c->Y * c->Z,
                        hipMemcpyDeviceToDevice, c->stream));
  return GSE_SUCCESS;
}

static gse_status_t step_outputs(const gse_step_desc_t* d, double fscale, GatherOut& o,
                                 int* flags) {
  if (d->virial) return fail(GSE_ERROR_NOT_SUPPORTED, "the virial is defined for GSE-k only");
  *flags = 0;
  if (d->forceX || d->forceY || d->forceZ) *flags |= GF_SPHERE;
  if (d->cubeX || d->cubeY || d->cubeZ) *flags |= GF_CUBE;
  if (d->potential) *flags |= GF_POT;
  if ((*flags & GF_SPHERE) && !(d->forceX && d->forceY && d->forceZ))
    return fail(GSE_ERROR_INVALID_VALUE, "forceX/Y/Z must be all set or all NULL");
  if ((*flags & GF_CUBE) && !(d->cubeX && d->cubeY && d->cubeZ))
    return fail(GSE_ERROR_INVALID_VALUE, "cubeX/Y/Z must be all set or all NULL");
  o.fx = d->forceX;
  o.fy = d->forceY;
  o.fz = d->forceZ;
  o.cx = d->cubeX;
  o.cy = d->cubeY;
  o.cz = d->cubeZ;
  o.pot = d->potential;
  o.fscale = o.pscale = o.vscale = fscale;
  return GSE_SUCCESS;
}

gse_status_t gse_gser_step(gse_gser_t h, int n, const void* xyzq, const gse_step_desc_t* d) {
  if (!h || !d) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  gse_gsek_s* c = h->core;
  GatherOut o;
  int flags = 0;
  GSECHK(step_outputs(d, c->ccelec, o, &flags));
  GSECHK(gse_gser_spread_charge1(h, n, xyzq));
  GSECHK(gse_gser_spread_charge2(h));
  GSECHK(gse_gser_solve_poisson(h));
  GSECHK(grid_energy(c, h->rho.p, c->phiM.p, 1.0, d->energy ? d->energy : c->dout.as<double>()));
  if (flags) GSECHK(c->eng.gather_binned(c->sigma1, c->rcut1(), flags, c->phiM.p, o, nullptr,
                                         c->stream));
  return GSE_SUCCESS;
}

// ------------------------------- GSE-u (CpuGSEu) ---------------------------
gse_status_t gse_gseu_create(gse_gseu_t* out, gse_dtype_t at, gse_dtype_t ct, int N, double b,
                             int X, int Y, int Z, double sigma, double kappa, double ls,
                             double ls1, double bx, double by, double bz, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (N <= 0 || N > 64) return fail(GSE_ERROR_INVALID_VALUE, "N must be in [1, 64] (CpuGSEu.cpp:17)");
  if (!(b > 1.0) || !std::isfinite(b)) return fail(GSE_ERROR_INVALID_VALUE, "b must be > 1 (CpuGSEu.cpp:18)");
  gse_gsek_t core = nullptr;
  GSECHK(gse_gsek_create(&core, at, ct, X, Y, Z, sigma, kappa, ls, ls1, bx, by, bz, ccelec));
  auto* h = new gse_gseu_s();
  h->core = core;
  h->N = N;
  h->b = b;
  gse_status_t st = GSE_SUCCESS;
  const size_t gb = tsize(core->ct_dbl) * (size_t)X * Y * Z;
  for (int j = 0; j < N && st == GSE_SUCCESS; j++) {
    // CpuGSEu.cpp:52-56
    const double sj = std::sqrt(sigma * sigma * std::pow(b, 2 * j) -
                                2 * core->sigma1 * core->sigma1);
    if (!(sj > 0) || !std::isfinite(sj)) {
      st = fail(GSE_ERROR_INVALID_VALUE, "sigma_%d is not positive", j);
      break;
    }
    h->sig.push_back(sj);
    h->fac.push_back(std::pow(2 * M_PI * sj * sj, 3.0 / 2.0) / std::pow(sj * sj, 1.0 / 2.0));
    auto* t3 = new gse_gseu_s::Taps3();
    h->taps.push_back(t3);
    st = make_ongrid_taps(core->ct_dbl, X, Y, Z, sj, ls1 * std::sqrt(2.0) * sj, bx, by, bz, t3->t,
                          t3->R);
    DevBuf* g = new DevBuf();
    h->terms.push_back(g);
    if (st == GSE_SUCCESS) st = g->ensure(gb);
  }
  if (st != GSE_SUCCESS) {
    gse_gseu_destroy(h);
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_gseu_destroy(gse_gseu_t h) {
  if (!h) return GSE_SUCCESS;
  gse_gsek_destroy(h->core);
  for (auto* t3 : h->taps) {
    for (int a = 0; a < 3; a++) t3->t[a].release();
    delete t3;
  }
  for (DevBuf* g : h->terms) {
    g->release();
    delete g;
  }
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_gseu_set_stream(gse_gseu_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return gse_gsek_set_stream(h->core, s);
}

gse_status_t gse_gseu_spread_charge1(gse_gseu_t h, int n, const void* xyzq) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return gse_gsek_spread_charge1(h->core, n, xyzq);
}

gse_status_t gse_gseu_solve_poisson(gse_gseu_t h) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  gse_gsek_s* c = h->core;
  const size_t np = (size_t)c->X * c->Y * c->Z;
  for (int j = 0; j < h->N; j++) {
    GSECHK(conv3(c->ct_dbl, c->X, c->Y, c->Z, h->taps[j]->R, h->taps[j]->t, c->rhoS.p,
                 h->terms[j]->p, c->tmp, c->tmp2, c->stream));
    HIPCHK(launch_grid_scale(c->ct_dbl, h->terms[j]->p, h->terms[j]->p, np, h->fac[j], c->stream));
  }
  // phiM = rhoG_{N-1} + ... + rhoG_0, added in that order (CpuGSEu.cpp:60-66)
  HIPCHK(hipMemcpyAsync(c->phiM.p, h->terms[h->N - 1]->p, tsize(c->ct_dbl) * np,
                        hipMemcpyDeviceToDevice, c->stream));
  for (int j = h->N - 2; j >= 0; j--)
    HIPCHK(launch_grid_add(c->ct_dbl, c->phiM.p, h->terms[j]->p, np, c->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_gseu_calculate_energy(gse_gseu_t h, double* energy) {
  if (!h || !energy) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  gse_gsek_s* c = h->core;
  GSECHK(grid_energy(c, c->rhoS.p, c->phiM.p, 2.0 * std::log(h->b), c->dout.as<double>()));
  return host_energy(c, energy);
}

gse_status_t gse_gseu_interpolate_force(gse_gseu_t h, int n, const void* xyzq, void* fx,
                                        void* fy, void* fz) {
  if (!h || (n > 0 && (!fx || !fy || !fz))) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  gse_gsek_s* c = h->core;
  GSECHK(c->eng.bin_particles(n, xyzq, c->bx, c->by, c->bz, c->sigma1, c->rcut1(), false,
                              c->stream));
  GatherOut o;
  o.fx = fx;
  o.fy = fy;
  o.fz = fz;
  o.fscale = 2.0 * std::log(h->b) * c->ccelec;  // CpuGSEu.cpp:78
  return c->eng.gather_binned(c->sigma1, c->rcut1(), GF_SPHERE, c->phiM.p, o, nullptr, c->stream);
}

gse_status_t gse_gseu_copy_grid(gse_gseu_t h, int which, void* dst) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  return gse_gsek_copy_grid(h->core, which, dst);
}

gse_status_t gse_gseu_step(gse_gseu_t h, int n, const void* xyzq, const gse_step_desc_t* d) {
  if (!h || !d) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  gse_gsek_s* c = h->core;
  GatherOut o;
  int flags = 0;
  GSECHK(step_outputs(d, 2.0 * std::log(h->b) * c->ccelec, o, &flags));
  GSECHK(gse_gseu_spread_charge1(h, n, xyzq));
  GSECHK(gse_gseu_solve_poisson(h));
  GSECHK(grid_energy(c, c->rhoS.p, c->phiM.p, 2.0 * std::log(h->b),
                     d->energy ? d->energy : c->dout.as<double>()));
  if (flags) GSECHK(c->eng.gather_binned(c->sigma1, c->rcut1(), flags, c->phiM.p, o, nullptr,
                                         c->stream));
  return GSE_SUCCESS;
}

// ------------------------------- batched -----------------------------------
}  // extern "C"

namespace {

// Lays out bins, slots, grids and spectra of all systems for the current
// launch shape and uploads the descriptors; builds the Green's tables and the
// batched FFT plans on the first call.
gse_status_t batch_layout(gse_batch_s* h, hipStream_t st) {
  HIPCHK(hipStreamSynchronize(st));  // buffers may be reallocated below
  h->release_graph();
  const int ns = h->nsys;
  const bool dbl = h->ct_dbl;
  const size_t ct = tsize(dbl);
  const bool acc32 = !dbl && g_tuning.spread_lds32;
  h->sys.assign(ns, BatchSys{});
  h->items.clear();
  h->cchunks.clear();
  h->kchunks.clear();
  h->groups.clear();
  long long bofs = 0, slot0 = 0, gofs = 0, kofs = 0, brick0 = 0;
  bool tiles = true;
  std::vector<int> nzs;
  for (int k = 0; k < ns; k++) {
    const int si = h->order[k];
    const int X = h->dims[3 * si], Y = h->dims[3 * si + 1], Z = h->dims[3 * si + 2];
    const double bx = h->boxes[3 * si], by = h->boxes[3 * si + 1], bz = h->boxes[3 * si + 2];
    BatchSys& S = h->sys[si];
    S.g = make_geom(X, Y, Z, bx, by, bz, h->rc1);
    S.b = choose_bins(S.g, acc32 ? 4 : 8, (size_t)g_tuning.brick_budget, h->bin_edge);
    if (S.b.direct)
      return fail(GSE_ERROR_NOT_SUPPORTED,
                  "system %d: the support (n = %d) does not fit the batched LDS path", si,
                  S.g.nx);
    S.n = h->counts[si];
    const long long c = 2LL * S.n / S.b.nbins + 32;
    S.cap_bin = (int)((c + 15) / 16 * 16);
    S.bofs = (int)bofs;
    S.slot0 = slot0;
    S.gofs = gofs;
    S.kofs = kofs;
    S.brick0 = brick0;
    brick0 += (long long)S.b.nbins * S.b.Lx * S.b.Ly * S.b.Lz;
    tiles = tiles && brick_tile_ok(S.g, S.b);
    bofs += S.b.nbins;
    slot0 += (long long)S.b.nbins * S.cap_bin;
    gofs += (long long)X * Y * Z;
    const long long K = (long long)(X / 2 + 1) * Y * Z;
    kofs += K;
    S.images = ((2 * S.g.nx + 1 + X - 1) / X) * ((2 * S.g.ny + 1 + Y - 1) / Y) *
               ((2 * S.g.nz + 1 + Z - 1) / Z);
    const double s1 = h->sigma1;
    S.inv2s2 = 1.0 / (2.0 * s1 * s1);
    S.pref = std::pow(2.0 * M_PI * s1 * s1, -1.5);
    S.pref_bound = S.pref * spread_weight_bound(S.g, s1, false);
    S.rc2 = h->rc1 * h->rc1;
    S.vcell = S.g.hx * S.g.hy * S.g.hz;
    S.prefF = std::pow(2.0 * M_PI, -1.5) * S.vcell * std::pow(s1, -5.0);
    S.prefP = S.pref * S.vcell;
    spread_lds_bytes(dbl, S.g, S.b, &S.spread_cap);
    gather_lds_bytes(dbl, S.g, S.b, &S.gather_cap);
    // k-pass chunks of 4096 elements, grid conversion chunks of 16384 points
    S.kblk0 = (int)h->kchunks.size();
    for (long long a = 0; a < K; a += 4096)
      h->kchunks.push_back(BatchChunk{si, (int)a, (int)std::min<long long>(4096, K - a), 0});
    S.nkblk = (int)h->kchunks.size() - S.kblk0;
    const long long P = (long long)X * Y * Z;
    for (long long a = 0; a < P; a += 16384)
      h->cchunks.push_back(BatchChunk{si, (int)a, (int)std::min<long long>(16384, P - a), 0});
    const int nz = 2 * S.g.nz + 1;
    if (std::find(nzs.begin(), nzs.end(), nz) == nzs.end()) nzs.push_back(nz);
  }
  h->total_bins = (int)bofs;
  h->total_slots = slot0;
  h->total_points = gofs;
  h->total_k = kofs;
  // spread / gather items, grouped by 2nz+1 (compile-time z extent)
  for (int nz : nzs) {
    BatchGroup gr;
    gr.nz = nz;
    gr.item0 = (int)h->items.size();
    for (int k = 0; k < ns; k++) {
      const int si = h->order[k];
      const BatchSys& S = h->sys[si];
      if (2 * S.g.nz + 1 != nz) continue;
      for (int b = 0; b < S.b.nbins; b++) h->items.push_back(make_int2(si, b));
      gr.lds_spread = std::max(gr.lds_spread, spread_lds_bytes(dbl, S.g, S.b, nullptr));
      gr.tile4 = gr.tile4 && acc32 && S.b.Lx % 4 == 0;
      gr.lds_gather = std::max(gr.lds_gather, gather_lds_bytes(dbl, S.g, S.b, nullptr));
      gr.pair = gr.pair && gather_pair_ok(dbl, h->at_dbl, S.g, S.b);
    }
    gr.nitems = (int)h->items.size() - gr.item0;
    const int want = std::min(gr.nitems, 2048);
    gr.grid = std::max(8, (want + 7) / 8 * 8);
    h->groups.push_back(gr);
  }
  // device buffers
  const int n1 = h->ntot > 0 ? h->ntot : 1;
  GSECHK(h->d_sys.ensure(sizeof(BatchSys) * ns));
  GSECHK(h->d_items.ensure(sizeof(int2) * std::max<size_t>(1, h->items.size())));
  GSECHK(h->d_cchunks.ensure(sizeof(BatchChunk) * std::max<size_t>(1, h->cchunks.size())));
  GSECHK(h->d_kchunks.ensure(sizeof(BatchChunk) * std::max<size_t>(1, h->kchunks.size())));
  GSECHK(h->counts_b.ensure(sizeof(int) * std::max(1, h->total_bins)));
  GSECHK(h->stats_b.ensure(sizeof(BinStats) * ns));
  GSECHK(h->ovf_b.ensure(sizeof(int) * n1));
  GSECHK(h->novf_b.ensure(sizeof(int)));
  GSECHK(h->perm_b.ensure(sizeof(int) * std::max<long long>(1, h->total_slots)));
  GSECHK(h->loc_b.ensure(sizeof(int) * std::max<long long>(1, h->total_slots)));
  GSECHK(h->sorted_b.ensure(4 * ct * std::max<long long>(1, h->total_slots)));
  GSECHK(h->grid64_b.ensure(8 * (size_t)h->total_points));
  GSECHK(h->rho_b.ensure(ct * (size_t)h->total_points));
  GSECHK(h->phi_b.ensure(ct * (size_t)h->total_points));
  GSECHK(h->spec_b.ensure(2 * ct * (size_t)h->total_k));
  GSECHK(h->table_b.ensure(ct * (size_t)h->total_k));
  GSECHK(h->part_b.ensure(sizeof(double) * std::max<size_t>(1, h->kchunks.size())));
  h->brick_flush = g_tuning.spread_flush && tiles;
  if (h->brick_flush) {
    GSECHK(h->bricks_b.ensure((acc32 ? 4 : 8) * (size_t)std::max<long long>(1, brick0)));
    GSECHK(h->bshift_b.ensure(sizeof(int) * std::max(1, h->total_bins)));
  }
  HIPCHK(hipMemcpy(h->d_sys.p, h->sys.data(), sizeof(BatchSys) * ns, hipMemcpyHostToDevice));
  if (!h->items.empty())
    HIPCHK(hipMemcpy(h->d_items.p, h->items.data(), sizeof(int2) * h->items.size(),
                     hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_cchunks.p, h->cchunks.data(), sizeof(BatchChunk) * h->cchunks.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_kchunks.p, h->kchunks.data(), sizeof(BatchChunk) * h->kchunks.size(),
                   hipMemcpyHostToDevice));
  std::vector<int> pofs(ns + 1, 0);
  for (int si = 0; si < ns; si++) pofs[si + 1] = pofs[si] + h->counts[si];
  GSECHK(h->d_pofs.ensure(sizeof(int) * (ns + 1)));
  HIPCHK(hipMemcpy(h->d_pofs.p, pofs.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
  BatchDev& d = h->dev;
  d.sys = h->d_sys.as<BatchSys>();
  d.nsys = ns;
  d.pofs = h->d_pofs.as<int>();
  d.ntot = h->ntot;
  d.counts = h->counts_b.as<int>();
  d.total_bins = h->total_bins;
  d.stats = h->stats_b.as<BinStats>();
  d.ovf = h->ovf_b.as<int>();
  d.novf = h->novf_b.as<int>();
  d.perm = h->perm_b.as<int>();
  d.loc = h->loc_b.as<int>();
  d.sorted = h->sorted_b.p;
  d.grid64 = h->grid64_b.as<unsigned long long>();
  d.bricks = h->brick_flush ? h->bricks_b.p : nullptr;
  d.bshift = h->brick_flush ? h->bshift_b.as<int>() : nullptr;
  HIPCHK(launch_fill_u64(d.grid64, (size_t)h->total_points, st));
  // Green's tables (G_GSE with sigma2) per system, batched FFT plans per grid size
  for (int k = 0; k < ns; k++) {
    const int si = h->order[k];
    const BatchSys& S = h->sys[si];
    HIPCHK(launch_greens_table(dbl, GSE_GREENS_GSE, S.g.X, S.g.Y, S.g.Z, h->boxes[3 * si],
                               h->boxes[3 * si + 1], h->boxes[3 * si + 2], h->sigma2,
                               (char*)h->table_b.p + ct * S.kofs, st));
  }
  if (h->ffts.empty()) {
    int k = 0;
    while (k < ns) {
      const int si = h->order[k];
      int e = k + 1;
      while (e < ns && h->dims[3 * h->order[e]] == h->dims[3 * si] &&
             h->dims[3 * h->order[e] + 1] == h->dims[3 * si + 1] &&
             h->dims[3 * h->order[e] + 2] == h->dims[3 * si + 2])
        e++;
      BatchFFT* f = new BatchFFT();
      h->ffts.push_back(f);
      GSECHK(f->fft.init(dbl, h->dims[3 * si], h->dims[3 * si + 1], h->dims[3 * si + 2], e - k));
      f->gofs = h->sys[si].gofs;
      f->kofs = h->sys[si].kofs;
      k = e;
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  return GSE_SUCCESS;
}

int batch_block(int want, bool dbl) { return want ? want : (dbl ? 256 : 512); }

// One batched step on stream st: every stage for all systems.
gse_status_t batch_run(gse_batch_s* h, hipStream_t st, const void* xyzq, void* fx, void* fy,
                       void* fz, void* pot, double* energies) {
  const bool dbl = h->ct_dbl;
  const size_t ct = tsize(dbl);
  const BatchDev& d = h->dev;
  HIPCHK(launch_batch_clear(d, st));
  HIPCHK(launch_batch_bin(dbl, d, xyzq, st));
  const int2* items = h->d_items.as<int2>();
  for (const BatchGroup& g : h->groups)
    HIPCHK(launch_batch_spread(dbl, g.nz, batch_block(h->spread_block, dbl), g.lds_spread, g.grid,
                               d, items + g.item0, g.nitems, st));
  HIPCHK(launch_batch_spread_overflow(dbl, d, xyzq, st));
  if (h->brick_flush) {
    const bool acc32 = !dbl && g_tuning.spread_lds32;
    for (const BatchGroup& g : h->groups)
      HIPCHK(launch_batch_reduce(dbl, acc32, g.tile4, d, items + g.item0, g.nitems, h->rho_b.p,
                                 st));
  } else {
    HIPCHK(launch_batch_convert(dbl, d, h->d_cchunks.as<BatchChunk>(), (int)h->cchunks.size(),
                                h->rho_b.p, st));
  }
  for (BatchFFT* f : h->ffts)
    GSECHK(f->fft.exec(true, (char*)h->rho_b.p + ct * f->gofs,
                       (char*)h->spec_b.p + 2 * ct * f->kofs, st));
  HIPCHK(launch_batch_kpass(dbl, d, h->d_kchunks.as<BatchChunk>(), (int)h->kchunks.size(),
                            h->spec_b.p, h->table_b.p, h->part_b.as<double>(), st));
  for (BatchFFT* f : h->ffts)
    GSECHK(f->fft.exec(false, (char*)h->spec_b.p + 2 * ct * f->kofs,
                       (char*)h->phi_b.p + ct * f->gofs, st));
  GatherOut o;
  o.fx = fx;
  o.fy = fy;
  o.fz = fz;
  o.pot = pot;
  o.fscale = o.pscale = o.vscale = h->ccelec;
  if (fx) {
    for (const BatchGroup& g : h->groups)
      HIPCHK(launch_batch_gather(dbl, h->at_dbl, g.nz, batch_block(h->gather_block, dbl),
                                 g.lds_gather, g.grid, d, items + g.item0, g.nitems, h->phi_b.p,
                                 o, g.pair, st));
    HIPCHK(launch_batch_gather_overflow(dbl, h->at_dbl, d, xyzq, h->phi_b.p, o, st));
  }
  if (energies)
    HIPCHK(launch_batch_energy(d, h->part_b.as<double>(), h->ccelec, energies, st));
  return GSE_SUCCESS;
}

gse_status_t batch_check_outputs(const gse_batch_s* h, const void* xyzq, void* fx, void* fy,
                                 void* fz) {
  if (h->ntot > 0 && !xyzq) return fail(GSE_ERROR_INVALID_VALUE, "xyzq is NULL");
  if ((fx || fy || fz) && !(fx && fy && fz))
    return fail(GSE_ERROR_INVALID_VALUE, "forceX/Y/Z must be all set or all NULL");
  return GSE_SUCCESS;
}

}  // namespace

extern "C" {

gse_status_t gse_batch_create(gse_batch_t* out, gse_dtype_t at, gse_dtype_t ct, int nsys,
                              const int* sizes, const double* boxes, const int* counts,
                              double sigma, double kappa, double ls, double ls1, double ccelec) {
  if (!out) return fail(GSE_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (!valid_pair(at, ct)) return fail(GSE_ERROR_INVALID_VALUE, "unsupported (AT, CT) pair");
  if (nsys <= 0 || !sizes || !boxes || !counts)
    return fail(GSE_ERROR_INVALID_VALUE, "need nsys > 0 and the sizes, boxes and counts arrays");
  if (!(kappa >= 0.0 && kappa <= 0.5)) return fail(GSE_ERROR_INVALID_VALUE, "kappa must be in [0, 0.5]");
  if (!(sigma > 0) || !(ls1 > 0)) return fail(GSE_ERROR_INVALID_VALUE, "sigma and lambdaSigma1 must be positive");
  long long ntot = 0;
  for (int k = 0; k < nsys; k++) {
    GSECHK(check_dims(sizes[3 * k], sizes[3 * k + 1], sizes[3 * k + 2]));
    GSECHK(check_box(boxes[3 * k], boxes[3 * k + 1], boxes[3 * k + 2]));
    if (counts[k] < 0) return fail(GSE_ERROR_INVALID_VALUE, "negative particle count");
    ntot += counts[k];
  }
  if (ntot > (1LL << 30)) return fail(GSE_ERROR_INVALID_VALUE, "too many particles");
  auto* h = new gse_batch_s();
  h->at_dbl = at == GSE_FLOAT64;
  h->ct_dbl = ct == GSE_FLOAT64;
  h->nsys = nsys;
  h->sigma = sigma;
  h->kappa = kappa;
  h->ls = ls;
  h->ls1 = ls1;
  h->ccelec = ccelec;
  h->sigma1 = sigma * std::sqrt(kappa);  // CpuGSE.cpp:42-43
  h->sigma2 = sigma * std::sqrt(1.0 - 2.0 * kappa);
  h->rc1 = ls1 * std::sqrt(2.0) * h->sigma1;
  h->dims.assign(sizes, sizes + 3 * nsys);
  h->boxes.assign(boxes, boxes + 3 * nsys);
  h->counts.assign(counts, counts + nsys);
  h->ntot = (int)ntot;
  h->order.resize(nsys);
  for (int k = 0; k < nsys; k++) h->order[k] = k;
  std::stable_sort(h->order.begin(), h->order.end(), [&](int a, int b) {
    for (int c = 0; c < 3; c++)
      if (h->dims[3 * a + c] != h->dims[3 * b + c]) return h->dims[3 * a + c] < h->dims[3 * b + c];
    return false;
  });
  gse_status_t st = batch_layout(h, nullptr);
  if (st != GSE_SUCCESS) {
    h->release();
    delete h;
    return st;
  }
  *out = h;
  return GSE_SUCCESS;
}

gse_status_t gse_batch_destroy(gse_batch_t h) {
  if (!h) return GSE_SUCCESS;
  h->release();
  delete h;
  return GSE_SUCCESS;
}

gse_status_t gse_batch_set_stream(gse_batch_t h, void* s) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  h->stream = (hipStream_t)s;
  return GSE_SUCCESS;
}

gse_status_t gse_batch_step(gse_batch_t h, const void* xyzq, void* fx, void* fy, void* fz,
                            void* pot, double* energies) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  GSECHK(batch_check_outputs(h, xyzq, fx, fy, fz));
  return batch_run(h, h->stream, xyzq, fx, fy, fz, pot, energies);
}

gse_status_t gse_batch_step_graph(gse_batch_t h, const void* xyzq, void* fx, void* fy, void* fz,
                                  void* pot, double* energies) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  GSECHK(batch_check_outputs(h, xyzq, fx, fy, fz));
  const void* ptrs[6] = {xyzq, fx, fy, fz, pot, energies};
  if (h->gexec && std::memcmp(ptrs, h->g_ptrs, sizeof(ptrs)) != 0) h->release_graph();
  if (!h->gexec) {
    if (!h->cap_stream) HIPCHK(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
    // the capture stream starts after the caller's pending work
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
    gse_status_t st = batch_run(h, h->cap_stream, xyzq, fx, fy, fz, pot, energies);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(h->cap_stream, &graph);
    if (st != GSE_SUCCESS) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return fail(GSE_ERROR_HIP, "stream capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      h->gexec = nullptr;
      return fail(GSE_ERROR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    std::memcpy(h->g_ptrs, ptrs, sizeof(ptrs));
  }
  HIPCHK(hipGraphLaunch(h->gexec, h->stream));
  return GSE_SUCCESS;
}

gse_status_t gse_batch_set_config(gse_batch_t h, int bin_edge, int spread_block,
                                  int gather_block) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (bin_edge < 1 || bin_edge > 64 || (spread_block != 0 && spread_block != 256 &&
                                        spread_block != 512) ||
      (gather_block != 0 && gather_block != 256 && gather_block != 512))
    return fail(GSE_ERROR_INVALID_VALUE, "bad launch shape");
  const int old[3] = {h->bin_edge, h->spread_block, h->gather_block};
  h->bin_edge = bin_edge;
  h->spread_block = spread_block;
  h->gather_block = gather_block;
  HIPCHK(hipStreamSynchronize(h->stream));
  gse_status_t st = batch_layout(h, h->stream);
  if (st != GSE_SUCCESS) {  // keep the previous, working shape
    h->bin_edge = old[0];
    h->spread_block = old[1];
    h->gather_block = old[2];
    (void)batch_layout(h, h->stream);
  }
  return st;
}

gse_status_t gse_batch_get_config(gse_batch_t h, int* config) {
  if (!h || !config) return fail(GSE_ERROR_INVALID_VALUE, "NULL argument");
  config[0] = h->bin_edge;
  config[1] = batch_block(h->spread_block, h->ct_dbl);
  config[2] = batch_block(h->gather_block, h->ct_dbl);
  return GSE_SUCCESS;
}

gse_status_t gse_batch_autotune(gse_batch_t h, const void* xyzq, int reps, double* best_ms) {
  if (!h) return fail(GSE_ERROR_INVALID_VALUE, "handle is NULL");
  if (h->ntot > 0 && !xyzq) return fail(GSE_ERROR_INVALID_VALUE, "xyzq is NULL");
  if (reps < 1) reps = 1;
  const size_t at = tsize(h->at_dbl);
  DevBuf f, e;
  GSECHK(f.ensure(3 * at * (size_t)(h->ntot > 0 ? h->ntot : 1)));
  GSECHK(e.ensure(sizeof(double) * h->nsys));
  char* fp = (char*)f.p;
  const size_t fb = at * (size_t)(h->ntot > 0 ? h->ntot : 1);
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  double best = 1e300;
  int bc[3] = {h->bin_edge, h->spread_block, h->gather_block};
  const int edges[3] = {6, 8, 10};
  const int blocks[2] = {256, 512};
  gse_status_t result = GSE_SUCCESS;
  for (int be : edges)
    for (int sb : blocks)
      for (int gb : blocks) {
        h->bin_edge = be;
        h->spread_block = sb;
        h->gather_block = gb;
        if (batch_layout(h, h->stream) != GSE_SUCCESS) continue;
        // one untimed step, then reps timed steps
        gse_status_t st = batch_run(h, h->stream, xyzq, fp, fp + fb, fp + 2 * fb, nullptr, e.as<double>());
        if (st != GSE_SUCCESS) continue;
        (void)hipEventRecord(t0, h->stream);
        for (int r = 0; r < reps && st == GSE_SUCCESS; r++)
          st = batch_run(h, h->stream, xyzq, fp, fp + fb, fp + 2 * fb, nullptr, e.as<double>());
        (void)hipEventRecord(t1, h->stream);
        if (hipEventSynchronize(t1) != hipSuccess || st != GSE_SUCCESS) continue;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (ms / reps < best) {
          best = ms / reps;
          bc[0] = be;
          bc[1] = sb;
          bc[2] = gb;
        }
      }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  h->bin_edge = bc[0];
  h->spread_block = bc[1];
  h->gather_block = bc[2];
  result = batch_layout(h, h->stream);
  f.release();
  e.release();
  if (best_ms) *best_ms = best < 1e300 ? best : -1.0;
  if (result == GSE_SUCCESS && !(best < 1e300))
    return fail(GSE_ERROR_HIP, "no launch shape ran");
  return result;
}

}  // extern "C"
