// Cell-binned Gaussian charge spreading with fixed-point accumulation
// (CpuGaussCharge::spreadChargeToGrid, CpuGaussCharge.cpp:235-384).
//
//   k_bin          one pass over the particles: cube origin
//                  o = wrap(round(x*inv_h) - n), bin of o, slot in the bin by
//                  atomic count; the particle, its original index and its
//                  cube origin within the bin are copied into the bin's fixed
//                  slots (or onto the overflow list); max|q| and rejected
//                  particles are counted.
//   k_spread       one workgroup per bin: the (B+2n)^3 brick lives in LDS
//                  (int64 at the global scale 2^e, or int32 at a per-bin
//                  scale 2^eb <= 2^e chosen from the bin's particle count).
//                  The bin's particles are staged into LDS, one wave per
//                  particle computes its 3(2n+1) factors, then one wave per
//                  particle walks its cube with lanes on (tx,ty) columns and
//                  adds round(q*pref*wx*wy*wz*2^eb) with LDS integer atomics.
//                  The brick is then either stored whole (with its shift
//                  e - eb) for the brick reduce, or added into the global
//                  int64 grid with 64-bit atomics.  kOverflowWG extra
//                  workgroups add the overflow particles into the global
//                  int64 grid at their home bin's scale.
//   k_brick_reduce[_tile|_tile4]  each output cell sums, exactly in int64,
//                  the stored bricks of the bins whose bricks cover it (each
//                  shifted to 2^e) plus the int64 overflow grid when the
//                  overflow list was used, then converts to T with 2^-e.
//   k_fixed_to_grid  int64 -> T with 2^-e, the int64 grid re-zeroed (atomic
//                  flush).
// Integer addition is associative, so the grid is bit-identical run to run
// whatever order particles and workgroups run in.  The scale 2^e follows from
// max|q|, the particle count and the support (fixed_exponent).
#include <math.h>

#include <algorithm>

#include "gse_device.h"
#include "gse_internal.h"

namespace gse {

template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

// XCD-aware workgroup -> bin mapping: workgroups are handed to the 8 XCDs
// round-robin, so give each XCD one contiguous range of bins (a slab of the
// grid); overlapping bricks then meet in the same L2.
__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
  const int chunk = (nwg + 7) >> 3;
  return (wg & 7) * chunk + (wg >> 3);
}

// The bin taken by spread/gather workgroup position p (order: NULL or the
// schedule of bin_order_eighth)
__device__ __forceinline__ int bin_at(const int* order, int p) { return order ? order[p] : p; }

__global__ void k_bin_clear(int* counts, int nbins, BinStats* stats) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nbins; i += gridDim.x * blockDim.x)
    counts[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stats->qmax_bits = 0ull;
    stats->invalid = 0;
    stats->overflow = 0;
    stats->spill = 0;
  }
}

template <typename CT>
__global__ void __launch_bounds__(256)
    k_bin(Geom g, BinShape b, int n, int cap_bin, const V4<CT>* __restrict__ xyzq,
          int* __restrict__ counts, BinStats* __restrict__ stats, int* __restrict__ ovf,
          int* __restrict__ perm, int* __restrict__ loc, V4<CT>* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double aq = 0.0;
  if (i < n) {
    const V4<CT> p = xyzq[i];
    // x*inv_h in FP64 as CpuGaussCharge.cpp:311-313 (CT=double); float
    // coordinates are promoted so cube selection matches the FP64 oracle
    const double fx = (double)p.x * g.ihx;
    const double fy = (double)p.y * g.ihy;
    const double fz = (double)p.z * g.ihz;
    const double lim = 1.0e9;
    const bool ok = fabs(fx) < lim && fabs(fy) < lim && fabs(fz) < lim &&
                    isfinite((double)p.w);  // NaN fails the comparisons
    if (ok) {
      const int ox = wrapi((int)round(fx) - g.nx, g.X);
      const int oy = wrapi((int)round(fy) - g.ny, g.Y);
      const int oz = wrapi((int)round(fz) - g.nz, g.Z);
      const int qx = ox / b.Bx, qy = oy / b.By, qz = oz / b.Bz;
      const int bin = qx + b.nbx * (qy + b.nby * qz);
      const int slot = atomicAdd(&counts[bin], 1);
      if (slot < cap_bin) {
        const size_t pos = (size_t)bin * cap_bin + slot;
        sorted[pos] = p;
        perm[pos] = i;
        loc[pos] = (ox - qx * b.Bx) | ((oy - qy * b.By) << 8) | ((oz - qz * b.Bz) << 16);
      } else {
        ovf[atomicAdd(&stats->overflow, 1)] = i;
      }
      aq = fabs((double)p.w);
    } else {
      atomicAdd(&stats->invalid, 1);
    }
  }
  // one max per workgroup reaches memory
  __shared__ double wmax[4];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) aq = fmax(aq, __shfl_xor(aq, m, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x >> 6] = aq;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    if (m > 0.0) atomicMax(&stats->qmax_bits, (unsigned long long)__double_as_longlong(m));
  }
}

// The same binning for grids of at most kBinAggMaxBins bins, with the slot
// claims aggregated per workgroup: a workgroup takes BT * kBinAggPPT
// consecutive particles, ranks them within their bins with
// LDS atomics, then claims each touched bin's slots with one global atomic
// (the particle's slot is that base plus its rank).  (BT threads, kBinAggPPT
// particles each.)  A random particle order
// touches each bin a few times per workgroup, so this issues several times
// fewer global atomics than one per particle.
constexpr int kBinAggMaxBins = 16384;
template <typename CT, int BT, int kBinAggPPT>
__global__ void __launch_bounds__(BT)
    k_bin_agg(Geom g, BinShape b, int n, int cap_bin, const V4<CT>* __restrict__ xyzq,
              int* __restrict__ counts, BinStats* __restrict__ stats, int* __restrict__ ovf,
              int* __restrict__ perm, int* __restrict__ loc, V4<CT>* __restrict__ sorted) {
  __shared__ int hist[kBinAggMaxBins];
  __shared__ double wmax[BT / 64];
  __shared__ int winv[BT / 64];
  for (int j = threadIdx.x; j < b.nbins; j += BT) hist[j] = 0;
  __syncthreads();
  const int i0 = blockIdx.x * (BT * kBinAggPPT);
  int bin_[kBinAggPPT], rank_[kBinAggPPT], loc_[kBinAggPPT];
  V4<CT> p_[kBinAggPPT];
  double aq = 0.0;
  int ninv = 0;
#pragma unroll
  for (int u = 0; u < kBinAggPPT; u++) {
    const int i = i0 + u * BT + (int)threadIdx.x;
    bin_[u] = -1;
    rank_[u] = 0;
    loc_[u] = 0;
    p_[u] = V4<CT>{0, 0, 0, 0};
    if (i < n) {
      const V4<CT> p = xyzq[i];
      p_[u] = p;
      // x*inv_h in FP64 as CpuGaussCharge.cpp:311-313, as k_bin
      const double fx = (double)p.x * g.ihx;
      const double fy = (double)p.y * g.ihy;
      const double fz = (double)p.z * g.ihz;
      const double lim = 1.0e9;
      const bool ok = fabs(fx) < lim && fabs(fy) < lim && fabs(fz) < lim && isfinite((double)p.w);
      if (ok) {
        const int ox = wrapi((int)round(fx) - g.nx, g.X);
        const int oy = wrapi((int)round(fy) - g.ny, g.Y);
        const int oz = wrapi((int)round(fz) - g.nz, g.Z);
        const int qx = ox / b.Bx, qy = oy / b.By, qz = oz / b.Bz;
        bin_[u] = qx + b.nbx * (qy + b.nby * qz);
        loc_[u] = (ox - qx * b.Bx) | ((oy - qy * b.By) << 8) | ((oz - qz * b.Bz) << 16);
        rank_[u] = atomicAdd(&hist[bin_[u]], 1);
        aq = fmax(aq, fabs((double)p.w));
      } else {
        ninv++;
      }
    }
  }
  __syncthreads();
  // one global claim per touched bin; hist[] becomes the bin's base slot
  for (int j = threadIdx.x; j < b.nbins; j += BT) {
    const int c = hist[j];
    if (c > 0) hist[j] = atomicAdd(&counts[j], c);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kBinAggPPT; u++) {
    if (bin_[u] < 0) continue;
    const int i = i0 + u * BT + (int)threadIdx.x;
    const int slot = hist[bin_[u]] + rank_[u];
    if (slot < cap_bin) {
      const size_t pos = (size_t)bin_[u] * cap_bin + slot;
      sorted[pos] = p_[u];
      perm[pos] = i;
      loc[pos] = loc_[u];
    } else {
      ovf[atomicAdd(&stats->overflow, 1)] = i;
    }
  }
  // one max and one invalid count per workgroup reach memory
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    aq = fmax(aq, __shfl_xor(aq, m, kWave));
    ninv += __shfl_xor(ninv, m, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wmax[threadIdx.x >> 6] = aq;
    winv[threadIdx.x >> 6] = ninv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    int inv = 0;
    for (int w = 0; w < BT / 64; w++) {
      m = fmax(m, wmax[w]);
      inv += winv[w];
    }
    if (m > 0.0) atomicMax(&stats->qmax_bits, (unsigned long long)__double_as_longlong(m));
    if (inv > 0) atomicAdd(&stats->invalid, inv);
  }
}

// Binning of grids with more than kBinAggMaxBins bins in two passes (a
// workgroup's randomly ordered charges touch most of the fine bins, so the
// claims of k_bin_agg would hardly aggregate, and their stores would scatter
// single charges over the whole slot array):
//   k_bin_coarse  super-bins of kSuperEdge^3 fine bins: each workgroup of
//                 BT x kBinSortPPT consecutive charges counts them per super-bin
//                 in LDS, claims one range per touched super-bin with one
//                 global atomic, and stores the charges and their indices
//                 there (runs of ~chunk / super-bins charges)
//   k_bin_fine    one workgroup per super-bin ranks its charges within their
//                 fine bins in LDS and stores them at their slots; it owns
//                 those bins, so it writes their counts directly, and it
//                 clears the super-bin's counter for the next call
// Charges past a super-bin's capacity are spilled to a list that the
// super-bin's k_bin_fine workgroup picks them from, so every bin's count is
// the number of charges located there (the fixed-point scale of the bin
// depends on it) and its first min(count, cap_bin) slots are filled; charges
// past a bin's cap_bin go to the overflow list.  The slots a charge gets differ from k_bin's only in order
// within a bin; the spread and gather do not depend on that order.
constexpr int kBinSortThreads = 1024;  // k_bin_fine
constexpr int kBinSortPPT = 4;          // charges per k_bin_coarse thread
constexpr int kSuperEdge = 4;                        // fine bins per super-bin edge
constexpr int kSuperBins = kSuperEdge * kSuperEdge * kSuperEdge;
constexpr int kSuperMax = 4096;                      // super-bins per grid

// Each super-bin's slots come in kStripes stripes; workgroup w claims in
// stripe w mod kStripes, so a counter takes the claims of 1/kStripes of the
// workgroups (claims on one address from every workgroup queue behind each
// other: 25.7 us for 245 workgroups x 343 super-bins at M = 216 with one
// stripe).
constexpr int kStripes = 8;

struct SuperShape {
  int nsx, nsy, nsz, nsuper, cap;  // cap: charge slots per stripe of a super-bin
};

static SuperShape super_shape(int n, const BinShape& b) {
  SuperShape q;
  q.nsx = (b.nbx + kSuperEdge - 1) / kSuperEdge;
  q.nsy = (b.nby + kSuperEdge - 1) / kSuperEdge;
  q.nsz = (b.nbz + kSuperEdge - 1) / kSuperEdge;
  q.nsuper = q.nsx * q.nsy * q.nsz;
  // twice the mean occupancy plus a margin, as cap_bin
  const long long c = 2LL * n / ((long long)(q.nsuper > 0 ? q.nsuper : 1) * kStripes) + 64;
  q.cap = (int)((c + 63) / 64 * 64);
  return q;
}

// scratch: [nsuper][kStripes] claim counters (zero between calls), then
// [nsuper][kStripes][cap] charges, their indices and their bin within the
// super-bin and in-bin cube origin (lc | lb << 24), and [n] spilled charges'
// indices (their count in BinStats::spill)
static size_t super_count_bytes(const SuperShape& q) {
  return ((size_t)q.nsuper * kStripes * 4 + 255) / 256 * 256;
}

size_t bin_sort_bytes(int n, const BinShape& b, bool dbl) {
  if (n <= 0 || b.direct || b.nbins <= kBinAggMaxBins) return 0;
  const SuperShape q = super_shape(n, b);
  if (q.nsuper > kSuperMax) return 0;
  const size_t slots = (size_t)q.nsuper * kStripes * q.cap;
  return super_count_bytes(q) + slots * (dbl ? 32 : 16) + slots * 8 + (size_t)n * 4;
}

// x mod m for x in [-m, 2m) without a division (any other x: the division)
__device__ __forceinline__ int wrap_near(int x, int m) {
  int r = x < 0 ? x + m : x;
  r = r >= m ? r - m : r;
  return (unsigned)r < (unsigned)m ? r : wrapi(x, m);
}
// x / d for 0 <= x < 2^12, 1 <= d <= 64 as a float product: the quotient of
// x + 1/2 lies at least 1/128 from an integer and the product's error is
// below 2^-10
__device__ __forceinline__ int div_small(int x, float inv_d) {
  return (int)(((float)x + 0.5f) * inv_d);
}

// bin and in-bin cube origin of a particle as k_bin computes them; false for
// a rejected particle
template <typename CT>
__device__ __forceinline__ bool bin_locate(const Geom& g, const BinShape& b, const V4<CT>& p,
                                           int& bin, int& lc, int& qx, int& qy, int& qz) {
  const double fx = (double)p.x * g.ihx;
  const double fy = (double)p.y * g.ihy;
  const double fz = (double)p.z * g.ihz;
  const double lim = 1.0e9;
  if (!(fabs(fx) < lim && fabs(fy) < lim && fabs(fz) < lim && isfinite((double)p.w)))
    return false;
  const int ox = wrap_near((int)round(fx) - g.nx, g.X);
  const int oy = wrap_near((int)round(fy) - g.ny, g.Y);
  const int oz = wrap_near((int)round(fz) - g.nz, g.Z);
  if (g.X < 4096 && g.Y < 4096 && g.Z < 4096) {
    qx = div_small(ox, 1.0f / (float)b.Bx);
    qy = div_small(oy, 1.0f / (float)b.By);
    qz = div_small(oz, 1.0f / (float)b.Bz);
  } else {
    qx = ox / b.Bx;
    qy = oy / b.By;
    qz = oz / b.Bz;
  }
  bin = qx + b.nbx * (qy + b.nby * qz);
  lc = (ox - qx * b.Bx) | ((oy - qy * b.By) << 8) | ((oz - qz * b.Bz) << 16);
  return true;
}

// Exclusive prefix sums out[0..n) of in[0..n) over the workgroup (BT
// threads, each summing a segment); returns the total.  wsum: BT/64 + 1
// entries of LDS; in and out must not overlap.  Starts and ends with every
// thread past a barrier.
template <int BT>
__device__ unsigned block_excl_scan(const unsigned* in, unsigned* out, int n, unsigned* wsum) {
  const int per = (n + BT - 1) / BT;
  const int j0 = (int)threadIdx.x * per;
  unsigned seg = 0;
  for (int k = 0; k < per; k++)
    if (j0 + k < n) seg += in[j0 + k];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned v = seg;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned r = 0;
    for (int k = 0; k < BT / 64; k++) {
      const unsigned t = wsum[k];
      wsum[k] = r;
      r += t;
    }
    wsum[BT / 64] = r;
  }
  __syncthreads();
  unsigned run = wsum[w] + v - seg;
  for (int k = 0; k < per; k++)
    if (j0 + k < n) {
      out[j0 + k] = run;
      run += in[j0 + k];
    }
  const unsigned total = wsum[BT / 64];
  __syncthreads();
  return total;
}

// LDS of k_bin_coarse: [nsuper] counts (then the claimed bases), [nsuper]
// prefix sums, then per staged charge its 4-vector, index, packed bin and
// in-bin origin, and super-bin
__host__ __device__ inline size_t coarse_lds_bytes(int nsuper, int chunk, int ct_bytes) {
  const size_t head = ((size_t)nsuper * 8 + 15) / 16 * 16;
  return head + (size_t)chunk * (4 * ct_bytes + 4 + 4 + 2);
}

template <typename CT, int BT>
__global__ void __launch_bounds__(BT)
    k_bin_coarse(Geom g, BinShape b, SuperShape q, int n, const V4<CT>* __restrict__ xyzq,
                 unsigned* __restrict__ scount, V4<CT>* __restrict__ sx, int* __restrict__ sidx,
                 int* __restrict__ sloc, int* __restrict__ spill, BinStats* __restrict__ stats) {
  constexpr int CH = BT * kBinSortPPT;
  extern __shared__ __align__(16) char csm[];
  unsigned* hs = (unsigned*)csm;
  unsigned* lpre = hs + q.nsuper;
  V4<CT>* stX = (V4<CT>*)(csm + ((size_t)q.nsuper * 8 + 15) / 16 * 16);
  int* stI = (int*)(stX + CH);
  int* stP = stI + CH;
  unsigned short* stS = (unsigned short*)(stP + CH);
  __shared__ double wmax[BT / 64];
  __shared__ int winv[BT / 64];
  __shared__ unsigned wsum[BT / 64 + 1];
  for (int j = threadIdx.x; j < q.nsuper; j += BT) hs[j] = 0u;
  __syncthreads();
  const int i0 = blockIdx.x * CH;
  double aq = 0.0;
  int ninv = 0;
  V4<CT> p_[kBinSortPPT];
  int sup_[kBinSortPPT], pk_[kBinSortPPT];
  unsigned rank_[kBinSortPPT];
#pragma unroll
  for (int u = 0; u < kBinSortPPT; u++) {
    const int i = i0 + u * BT + (int)threadIdx.x;
    sup_[u] = -1;
    if (i >= n) continue;
    p_[u] = xyzq[i];
    int bin, lc, qx, qy, qz;
    if (bin_locate<CT>(g, b, p_[u], bin, lc, qx, qy, qz)) {
      sup_[u] = qx / kSuperEdge + q.nsx * (qy / kSuperEdge + q.nsy * (qz / kSuperEdge));
      pk_[u] = lc | ((qx % kSuperEdge) + kSuperEdge * ((qy % kSuperEdge) +
                                                       kSuperEdge * (qz % kSuperEdge))) << 24;
      rank_[u] = atomicAdd(&hs[sup_[u]], 1u);
      aq = fmax(aq, fabs((double)p_[u].w));
    } else {
      ninv++;
    }
  }
  __syncthreads();
  // the charges staged in LDS in super-bin order, so that each super-bin's
  // run reaches memory as consecutive lanes' stores
  const int nloc = (int)block_excl_scan<BT>(hs, lpre, q.nsuper, wsum);
  // one global claim per touched super-bin in this workgroup's stripe; hs[]
  // becomes the run's base
  const int stripe = blockIdx.x % kStripes;
  for (int j = threadIdx.x; j < q.nsuper; j += BT) {
    const unsigned c = hs[j];
    if (c > 0u) hs[j] = atomicAdd(&scount[j * kStripes + stripe], c);
  }
#pragma unroll
  for (int u = 0; u < kBinSortPPT; u++) {
    if (sup_[u] < 0) continue;
    const int e = (int)(lpre[sup_[u]] + rank_[u]);
    stX[e] = p_[u];
    stI[e] = i0 + u * BT + (int)threadIdx.x;
    stP[e] = pk_[u];
    stS[e] = (unsigned short)sup_[u];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nloc; e += BT) {
    const int sp = stS[e];
    const unsigned slot = hs[sp] + (unsigned)(e - (int)lpre[sp]);
    if (slot < (unsigned)q.cap) {
      const size_t pos = ((size_t)sp * kStripes + stripe) * q.cap + slot;
      sx[pos] = stX[e];
      sidx[pos] = stI[e];
      sloc[pos] = stP[e];
    } else {
      spill[atomicAdd(&stats->spill, 1)] = stI[e];
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    aq = fmax(aq, __shfl_xor(aq, m, kWave));
    ninv += __shfl_xor(ninv, m, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wmax[threadIdx.x >> 6] = aq;
    winv[threadIdx.x >> 6] = ninv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    int inv = 0;
    for (int w = 0; w < BT / 64; w++) {
      m = fmax(m, wmax[w]);
      inv += winv[w];
    }
    if (m > 0.0) atomicMax(&stats->qmax_bits, (unsigned long long)__double_as_longlong(m));
    if (inv > 0) atomicAdd(&stats->invalid, inv);
  }
}

// one charge into its bin's next slot (lb: its bin within the super-bin)
template <typename CT>
__device__ __forceinline__ void bin_fine_place(const BinShape& b, int ux, int uy, int uz, int lb,
                                               int lc, const V4<CT>& p, int i, int cap_bin,
                                               int* cnt, BinStats* stats, int* ovf, int* perm,
                                               int* loc, V4<CT>* sorted) {
  const int qx = kSuperEdge * ux + lb % kSuperEdge;
  const int qy = kSuperEdge * uy + (lb / kSuperEdge) % kSuperEdge;
  const int qz = kSuperEdge * uz + lb / (kSuperEdge * kSuperEdge);
  const int bin = qx + b.nbx * (qy + b.nby * qz);
  const int slot = atomicAdd(&cnt[lb], 1);
  if (slot < cap_bin) {
    const size_t pos = (size_t)bin * cap_bin + slot;
    sorted[pos] = p;
    perm[pos] = i;
    loc[pos] = lc;
  } else {
    ovf[atomicAdd(&stats->overflow, 1)] = i;
  }
}

constexpr int kBinFineLoads = 4;  // charges loaded before they are placed

// flat index k of a super-bin's charges -> offset in its striped slots
// (pre: first flat index of each stripe)
__device__ __forceinline__ size_t stripe_offset(int k, const int* pre, int cap) {
  int j = 0;
#pragma unroll
  for (int t = 1; t < kStripes; t++) j += k >= pre[t] ? 1 : 0;
  return (size_t)j * cap + (size_t)(k - pre[j]);
}

// (the fine pass stores each charge at its slot directly: staging its
// charges in LDS in bin order first, for runs of consecutive stores,
// measured slower, 23.6 against 18.1 us at M = 216)
template <typename CT>
__global__ void __launch_bounds__(kBinSortThreads)
    k_bin_fine(Geom g, BinShape b, SuperShape q, int cap_bin, const V4<CT>* __restrict__ xyzq,
               unsigned* __restrict__ scount, const V4<CT>* __restrict__ sx,
               const int* __restrict__ sidx, const int* __restrict__ sloc,
               const int* __restrict__ spill, BinStats* __restrict__ stats,
               int* __restrict__ ovf, int* __restrict__ counts, int* __restrict__ perm,
               int* __restrict__ loc, V4<CT>* __restrict__ sorted) {
  __shared__ int cnt[kSuperBins];
  __shared__ int pre[kStripes + 1];  // first charge of each stripe
  const int s = blockIdx.x;
  const int ux = s % q.nsx, uy = (s / q.nsx) % q.nsy, uz = s / (q.nsx * q.nsy);
  if (threadIdx.x < kSuperBins) cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    int run = 0;
    for (int j = 0; j < kStripes; j++) {
      pre[j] = run;
      run += (int)min(scount[s * kStripes + j], (unsigned)q.cap);
    }
    pre[kStripes] = run;
  }
  __syncthreads();
  const int ns = pre[kStripes];
  const int nsp = stats->spill;
  const size_t base = (size_t)s * kStripes * q.cap;
  for (int k0 = 0; k0 < ns; k0 += kBinFineLoads * kBinSortThreads) {
    V4<CT> p[kBinFineLoads];
    int idx[kBinFineLoads], pk[kBinFineLoads];
#pragma unroll
    for (int u = 0; u < kBinFineLoads; u++) {
      const int k = min(k0 + u * kBinSortThreads + (int)threadIdx.x, ns - 1);
      const size_t at = base + stripe_offset(k, pre, q.cap);
      p[u] = sx[at];
      idx[u] = sidx[at];
      pk[u] = sloc[at];
    }
#pragma unroll
    for (int u = 0; u < kBinFineLoads; u++) {
      const int k = k0 + u * kBinSortThreads + (int)threadIdx.x;
      if (k < ns)
        bin_fine_place<CT>(b, ux, uy, uz, pk[u] >> 24, pk[u] & 0xffffff, p[u], idx[u], cap_bin,
                           cnt, stats, ovf, perm, loc, sorted);
    }
  }
  // charges that found their super-bin full (none unless it holds more
  // than twice the mean)
  for (int k = threadIdx.x; k < nsp; k += kBinSortThreads) {
    const int i = spill[k];
    const V4<CT> p = xyzq[i];
    int bin, lc, qx, qy, qz;
    bin_locate<CT>(g, b, p, bin, lc, qx, qy, qz);  // accepted by k_bin_coarse
    if (qx / kSuperEdge == ux && qy / kSuperEdge == uy && qz / kSuperEdge == uz)
      bin_fine_place<CT>(b, ux, uy, uz,
                         (qx % kSuperEdge) +
                             kSuperEdge * ((qy % kSuperEdge) + kSuperEdge * (qz % kSuperEdge)),
                         lc, p, i, cap_bin, cnt, stats, ovf, perm, loc, sorted);
  }
  __syncthreads();
  if (threadIdx.x < kSuperBins) {
    const int lx = threadIdx.x % kSuperEdge, ly = (threadIdx.x / kSuperEdge) % kSuperEdge,
              lz = threadIdx.x / (kSuperEdge * kSuperEdge);
    const int qx = kSuperEdge * ux + lx, qy = kSuperEdge * uy + ly, qz = kSuperEdge * uz + lz;
    if (qx < b.nbx && qy < b.nby && qz < b.nbz)
      counts[qx + b.nbx * (qy + b.nby * qz)] = cnt[threadIdx.x];
  }
  if (threadIdx.x < kStripes) scount[s * kStripes + threadIdx.x] = 0u;  // for the next call
}

struct LaneRole {
  int axis;  // 0,1,2 or -1
  int j;
};

// Weight of (axis, j): exp(-t*t*inv_2sigmasq) with t = (c*h - x) + j*h
// (CpuGaussCharge.cpp:314-329); offsets in FP64, the exponential in CT.
template <typename CT>
__device__ __forceinline__ CT lane_weight(const LaneRole& r, int cx, int cy, int cz, double x,
                                          double y, double z, const Geom& g, double inv2s2) {
  if (r.axis < 0) return (CT)0;
  double c, h, xv;
  if (r.axis == 0) {
    c = (double)cx; h = g.hx; xv = x;
  } else if (r.axis == 1) {
    c = (double)cy; h = g.hy; xv = y;
  } else {
    c = (double)cz; h = g.hz; xv = z;
  }
  const double d = sub_rn(mul_rn(c, h), xv);
  const double t = add_rn(d, mul_rn((double)r.j, h));
  const CT tc = (CT)t;
  return gexp(mul_rn(mul_rn(-tc, tc), (CT)inv2s2));
}

constexpr int kMaxK = 5;  // ceil((2n+1)^2/64) for n <= 8

// Adds one particle's cube straight into the global int64 grid (overflow
// particles and supports too wide for the LDS path); one wave per particle.
// S is the particle's scale; each rounded contribution is shifted left by
// `shift` (the int32 path rounds at its home bin's scale 2^eb and moves to
// the global 2^e exactly, so a particle adds the same integers whether it
// went through its bin's brick or the overflow list).
template <typename CT>
__device__ void spread_one_global(const Geom& g, double inv2s2, double pref, double S,
                                  int normalize, double pref1, const V4<CT>& p,
                                  unsigned long long* __restrict__ grid64, int shift = 0) {
  const int lane = threadIdx.x & 63;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
  const int cx = (int)round(x * g.ihx) - g.nx;
  const int cy = (int)round(y * g.ihy) - g.ny;
  const int cz = (int)round(z * g.ihz) - g.nz;
  const double qps = ((double)p.w * pref) * S;
  CT sx = 1, sy = 1, sz = 1;
  if (normalize) {  // per-axis sums in j order (CpuGaussCharge.cpp:330-362)
    LaneRole r;
    sx = sy = sz = 0;
    r.axis = 0;
    for (int j = 0; j < NX; j++) { r.j = j; sx += lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2); }
    r.axis = 1;
    for (int j = 0; j < NY; j++) { r.j = j; sy += lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2); }
    r.axis = 2;
    for (int j = 0; j < NZ; j++) { r.j = j; sz += lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2); }
  }
  const int wx0 = wrapi(cx, g.X), wy0 = wrapi(cy, g.Y), wz0 = wrapi(cz, g.Z);
  for (int idx = lane; idx < NX * NY * NZ; idx += 64) {
    const int tx = idx % NX, t2 = idx / NX, ty = t2 % NY, tz = t2 / NY;
    LaneRole r;
    r.axis = 0; r.j = tx;
    CT wx = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
    r.axis = 1; r.j = ty;
    CT wy = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
    r.axis = 2; r.j = tz;
    CT wzv = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
    if (normalize) {
      wx = wx * ((CT)pref1 / sx);
      wy = wy * ((CT)pref1 / sy);
      wzv = wzv * ((CT)pref1 / sz);
    }
    const double qxy = (qps * (double)wx) * (double)wy;
    const int gx = (wx0 + tx) % g.X, gy = (wy0 + ty) % g.Y, gz = (wz0 + tz) % g.Z;
    atomicAdd(&grid64[gx + (size_t)g.X * (gy + (size_t)g.Y * gz)],
              (unsigned long long)(fixed_round_prod(qxy, (double)wzv) << shift));
  }
}

struct SpreadArgs {
  Geom g;
  BinShape b;
  double inv2s2, pref, pref1;
  int normalize;
  int cap;      // particles staged in LDS per chunk
  int cap_bin;  // slots per bin
  int n;
  int images;
  double pref_bound;
  int nwg_bins;
  const int* order;  // bin of each workgroup position (NULL: the position itself)
};

// LDS layout of k_spread: int64 brick at 0, then per staged particle (cap
// of them) the 4-vector (CT, 32-byte aligned block), its 3(2n+1) factors
// (CT, as computed), q*pref*2^e (double) and its local brick offset (int).
struct SpreadLds {
  size_t oX, oW, oQ, oB, total;
};
__host__ __device__ inline SpreadLds spread_lds(int vol, int cap, int nw, int ct_bytes) {
  SpreadLds l;
  size_t o = ((size_t)vol * 8 + 31) & ~(size_t)31;
  l.oX = o;
  o += (size_t)cap * 4 * ct_bytes;
  l.oW = o;
  o += ((size_t)cap * nw * ct_bytes + 7) & ~(size_t)7;
  l.oQ = o;
  o += (size_t)cap * 8;
  l.oB = o;
  o += (size_t)cap * 4;
  l.total = o;
  return l;
}

// Bin of a particle: the bin of its wrapped cube origin (as k_bin files it).
template <typename CT>
__device__ __forceinline__ int home_bin(const Geom& g, const BinShape& b, const V4<CT>& p) {
  const int ox = wrapi((int)round((double)p.x * g.ihx) - g.nx, g.X);
  const int oy = wrapi((int)round((double)p.y * g.ihy) - g.ny, g.Y);
  const int oz = wrapi((int)round((double)p.z * g.ihz) - g.nz, g.Z);
  return ox / b.Bx + b.nbx * (oy / b.By + b.nby * (oz / b.Bz));
}

// One overflow particle (one wave) straight into the int64 grid, at the scale
// its bin's brick would have used.
template <typename CT, typename ACC>
__device__ __forceinline__ void spread_overflow(const SpreadArgs& a, int eg,
                                                unsigned long long qbits,
                                                const int* __restrict__ counts, const V4<CT>& p,
                                                unsigned long long* __restrict__ grid64) {
  int shift = 0;
  if (sizeof(ACC) == 4) {
    const int eb = fixed_exponent_bin(qbits, counts[home_bin<CT>(a.g, a.b, p)], a.pref_bound);
    shift = eg > eb ? eg - eb : 0;
  }
  spread_one_global<CT>(a.g, a.inv2s2, a.pref, ldexp(1.0, eg - shift), a.normalize, a.pref1, p,
                        grid64, shift);
}

// ACC: LDS accumulator, unsigned long long (int64 at the global scale 2^e) or
// unsigned (int32 at a per-bin scale 2^eb, eb <= e, from the bin's particle
// count, shifted exactly to 2^e when the brick is flushed).
//
// One bin of one system: the bin's cnt staged particles start at `sorted`
// (slot 0 of the bin), the system's int64 grid at `grid64`; eg is the
// system's exponent, qbits its max |q| and cnt_all the bin's particle count
// including overflow (it sets the int32 scale).  The finished brick is either
// added into grid64 with 64-bit atomics (brick_out NULL) or stored as it is
// into brick_out, with its shift to the global scale in shift_out[bin], for
// k_brick_reduce.  Ends with a barrier, so a workgroup may call it for
// several bins in a row.
template <typename CT, int NZC, int BS, typename ACC, int PL = 0>
__device__ __forceinline__ void spread_bin(const SpreadArgs& a, int bin, int cnt, int cnt_all,
                                           int eg, unsigned long long qbits,
                                           const V4<CT>* __restrict__ sorted,
                                           const int* __restrict__ loc,
                                           unsigned long long* __restrict__ grid64,
                                           unsigned char* smem, ACC* __restrict__ brick_out,
                                           int* __restrict__ shift_out) {
  ACC* const brick = (ACC*)smem;
  const Geom& g = a.g;
  const BinShape& b = a.b;
  const double S = ldexp(1.0, eg);
  // int32 bricks: per-bin exponent so that the bin's brick sums stay below 2^30
  int shift = 0;
  double Sb = S;
  if (sizeof(ACC) == 4) {
    const int eb = fixed_exponent_bin(qbits, cnt_all, a.pref_bound);
    shift = eg > eb ? eg - eb : 0;
    Sb = ldexp(S, -shift);
  }
  const int bxi = bin % b.nbx, byi = (bin / b.nbx) % b.nby, bzi = bin / (b.nbx * b.nby);
  const int Lx = b.Lx, Ly = b.Ly, Lz = b.Lz, LxLy = b.Lx * b.Ly;
  const int vol = LxLy * Lz;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1;
  const int NZ = NZC > 0 ? NZC : 2 * g.nz + 1;
  const int NW = NX + NY + NZ;
  const int cap = a.cap;
  const SpreadLds lay = spread_lds(vol * (int)sizeof(ACC) / 8 + 1, cap, NW, sizeof(CT));
  char* lds = (char*)brick;
  V4<CT>* sX = (V4<CT>*)(lds + lay.oX);  // [cap]
  CT* sW = (CT*)(lds + lay.oW);  // [cap][NW] factors (values of type CT)
  double* sQ = (double*)(lds + lay.oQ);  // [cap]
  int* sB = (int*)(lds + lay.oB);        // [cap]
  // the first chunk's particle (cap <= 256 <= BS: at most one per thread) is
  // loaded before the brick is cleared, so its latency overlaps the clear
  // (the index is clamped into the bin's slots, so the load needs no branch
  // and the values stay in registers; slots past cnt are never used)
  const int pslot = min((int)threadIdx.x, a.cap_bin - 1);
  const V4<CT> pre = sorted[pslot];
  const int pre_l = loc[pslot];
  for (int j = threadIdx.x; j < vol; j += BS) brick[j] = (ACC)0;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int nwv = BS / 64;
  const int NXY = NX * NY;
  const int KI = (NXY + 63) >> 6;
  LaneRole role;
  if (lane < NX) {
    role.axis = 0; role.j = lane;
  } else if (lane < NX + NY) {
    role.axis = 1; role.j = lane - NX;
  } else if (lane < NW) {
    role.axis = 2; role.j = lane - NX - NY;
  } else {
    role.axis = -1; role.j = 0;
  }
  // this lane's factor axis: spacing, inverse spacing and support half-width
  // (selected once, as values, so no field of `a` is indexed at run time)
  double lh = 0.0, lih = 0.0;
  int ln = 0;
  {
    const double hx = g.hx, hy = g.hy, hz = g.hz, ihx = g.ihx, ihy = g.ihy, ihz = g.ihz;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (role.axis == 0) {
      lh = hx; lih = ihx; ln = nx;
    } else if (role.axis == 1) {
      lh = hy; lih = ihy; ln = ny;
    } else {
      lh = hz; lih = ihz; ln = nz;
    }
  }
  const double lane_j = (double)role.j;
  const CT inv2s2_c = (CT)a.inv2s2;
  // column k of this lane: factor indices k_tx, NX + k_ty and brick offset
  // k_off (-1: no column)
  int k_tx[kMaxK], k_ty[kMaxK], k_off[kMaxK];
#pragma unroll
  for (int k = 0; k < kMaxK; k++) {
    const int idx = lane + 64 * k;
    const bool ok = (k < KI) && idx < NXY;
    const int ty = ok ? idx / NX : 0;
    k_tx[k] = ok ? idx - ty * NX : 0;
    k_ty[k] = NX + ty;
    k_off[k] = ok ? k_tx[k] + Lx * ty : -1;
  }

  for (int c0 = 0; c0 < cnt; c0 += cap) {
    const int na = min(cap, cnt - c0);
    // phase 0: the chunk's particles into LDS in one coalesced pass, with
    // their brick offsets (from the cube origin k_bin stored) and scaled charges
    if ((int)threadIdx.x < na) {
      CT px = pre.x, py = pre.y, pz = pre.z, pw = pre.w;
      int l = pre_l;
      if (c0 != 0) {
        const V4<CT> q = sorted[c0 + threadIdx.x];
        px = q.x, py = q.y, pz = q.z, pw = q.w;
        l = loc[c0 + threadIdx.x];
      }
      const V4<CT> p{px, py, pz, pw};
      sX[threadIdx.x] = p;
      sB[threadIdx.x] = (l & 255) + Lx * (((l >> 8) & 255) + Ly * (l >> 16));
      sQ[threadIdx.x] = ((double)p.w * a.pref) * Sb;
    }
    __syncthreads();
    // phase 1: one wave per particle, lane j computes factor j of its axis
    for (int ai = wid; ai < na; ai += nwv) {
      if (role.axis >= 0) {
        const V4<CT> p = sX[ai];
        const double xv = role.axis == 0 ? (double)p.x : role.axis == 1 ? (double)p.y : (double)p.z;
        const int c = (int)round(xv * lih) - ln;
        // t = (c*h - x) + j*h, rounded as lane_weight does (CpuGaussCharge.cpp:314-329)
        const double t = add_rn(sub_rn(mul_rn((double)c, lh), xv), mul_rn(lane_j, lh));
        const CT tc = (CT)t;
        sW[(size_t)ai * NW + lane] = gexp(mul_rn(mul_rn(-tc, tc), inv2s2_c));
      }
    }
    __syncthreads();
    if (a.normalize) {  // CpuGaussCharge.cpp:330-362: per-axis sums in j order, in CT
      for (int e = threadIdx.x; e < na * 3; e += BS) {
        const int ai = e / 3, ax = e - 3 * ai;
        const int j0 = ax == 0 ? 0 : ax == 1 ? NX : NX + NY;
        const int nj = ax == 0 ? NX : ax == 1 ? NY : NZ;
        CT* w = sW + (size_t)ai * NW + j0;
        CT sum = 0;
        for (int j = 0; j < nj; j++) sum += w[j];
        const CT nrm = (CT)a.pref1 / sum;
        for (int j = 0; j < nj; j++) w[j] = w[j] * nrm;
      }
      __syncthreads();
    }
    // phase 2: one wave per particle, lanes on (tx,ty) columns, walk tz
    for (int ai = wid; ai < na; ai += nwv) {
      const CT* w = sW + (size_t)ai * NW;
      const int base = sB[ai];
      const double qps = sQ[ai];
      double wz[NZC > 0 ? NZC : 1];
      if (NZC > 0) {
#pragma unroll
        for (int t = 0; t < (NZC > 0 ? NZC : 1); t++) wz[t] = (double)w[NX + NY + t];
      }
#pragma unroll
      for (int k = 0; k < kMaxK; k++) {
        if (k >= KI) break;
        if (k_off[k] >= 0) {
          // ((q pref 2^e) wx) wy: the reference's product order, scaled exactly
          const double qxy = (qps * (double)w[k_tx[k]]) * (double)w[k_ty[k]];
          const int a0 = base + k_off[k];
          if (NZC > 0) {
            // PL > 0: the plane stride is a compile-time LDS offset
            const int pl = PL > 0 ? PL : LxLy;
#pragma unroll
            for (int t = 0; t < (NZC > 0 ? NZC : 1); t++)
              atomicAdd(&brick[a0 + t * pl], (ACC)fixed_round_prod(qxy, wz[t]));
          } else {
            for (int t = 0; t < NZ; t++)
              atomicAdd(&brick[a0 + t * LxLy], (ACC)fixed_round_prod(qxy, w[NX + NY + t]));
          }
        }
      }
    }
    __syncthreads();
  }

  if (brick_out) {  // plain coalesced stores; k_brick_reduce sums the bricks
    if (sizeof(ACC) == 4 && (vol & 3) == 0) {  // 16-byte pieces (the brick is 32-byte aligned)
      const int vq = vol >> 2;
      for (int j = threadIdx.x; j < vq; j += BS) ((uint4*)brick_out)[j] = ((const uint4*)brick)[j];
    } else {
      for (int j = threadIdx.x; j < vol; j += BS) brick_out[j] = brick[j];
    }
    if (threadIdx.x == 0) shift_out[bin] = shift;
    __syncthreads();
    return;
  }
  // flush: rows of Lx lanes (rows_per = 64/Lx per wave), no per-element
  // division; (ly, lz) advance by the constant row stride
  const int X = g.X, Y = g.Y, Z = g.Z;
  const int rows_per = 64 / Lx;
  const int rl = lane / Lx;
  const int lx = lane - rl * Lx;
  int gx = bxi * b.Bx + lx;
  while (gx >= X) gx -= X;
  const int rows = Ly * Lz;
  const int rstep = nwv * rows_per;
  int r = wid * rows_per + rl;
  int ly = r % Ly, lz = r / Ly;
  for (; r < rows; r += rstep) {
    if (rl < rows_per) {
      const ACC raw = brick[lx + Lx * (ly + Ly * lz)];
      // sign-extend int32 bricks and move them to the global scale 2^e
      const unsigned long long v =
          sizeof(ACC) == 4 ? (unsigned long long)((long long)(int)raw << shift)
                           : (unsigned long long)raw;
      if (v != 0ull) {
        int gy = byi * b.By + ly, gz = bzi * b.Bz + lz;
        while (gy >= Y) gy -= Y;
        while (gz >= Z) gz -= Z;
        atomicAdd(&grid64[gx + (size_t)X * (gy + (size_t)Y * gz)], v);
      }
    }
    ly += rstep;
    while (ly >= Ly) {
      ly -= Ly;
      lz++;
    }
  }
  __syncthreads();
}

// 512-thread workgroups: three per CU (6 waves per SIMD, at most 80 VGPRs),
// so one workgroup's staging overlaps two others' work
template <typename CT, int NZC, int BS, typename ACC, int PL = 0>
__global__ void __launch_bounds__(BS, BS == 512 ? 6 : 1)
    k_spread(SpreadArgs a, const V4<CT>* __restrict__ xyzq, const V4<CT>* __restrict__ sorted,
             const int* __restrict__ counts, int* __restrict__ counts_other,
             const BinStats* __restrict__ stats, BinStats* __restrict__ stats_other,
             const int* __restrict__ ovf, unsigned long long* __restrict__ grid64,
             ACC* __restrict__ bricks, int* __restrict__ bshift, const int* __restrict__ loc) {
  extern __shared__ __align__(32) unsigned char spread_smem[];
  const BinStats st = *stats;
  const int eg = fixed_exponent(st.qmax_bits, a.n - st.invalid, a.images, a.pref_bound);
  if ((int)blockIdx.x >= a.nwg_bins) {  // overflow particles, global atomics
    if (blockIdx.x == (unsigned)a.nwg_bins && threadIdx.x == 0) {
      stats_other->qmax_bits = 0ull;
      stats_other->invalid = 0;
      stats_other->overflow = 0;
      stats_other->spill = 0;
    }
    const int wg = blockIdx.x - a.nwg_bins;
    const int gw = wg * (BS / 64) + (threadIdx.x >> 6);
    for (int k = __builtin_amdgcn_readfirstlane(gw); k < st.overflow; k += kOverflowWG * (BS / 64))
      spread_overflow<CT, ACC>(a, eg, st.qmax_bits, counts, xyzq[ovf[k]], grid64);
    return;
  }
  const int pos = xcd_remap(blockIdx.x, a.nwg_bins);
  if (pos >= a.b.nbins) return;
  const int bin = bin_at(a.order, pos);
  const int call = counts[bin];
  const int cnt = min(call, a.cap_bin);
  if (threadIdx.x == 0) {
    counts_other[bin] = 0;
    if (bricks && cnt == 0) bshift[bin] = -1;  // no brick this step
  }
  if (cnt == 0) return;
  const size_t vol = (size_t)a.b.Lx * a.b.Ly * a.b.Lz;
  spread_bin<CT, NZC, BS, ACC, PL>(a, bin, cnt, call, eg, st.qmax_bits,
                               sorted + (size_t)bin * a.cap_bin, loc + (size_t)bin * a.cap_bin,
                               grid64, spread_smem,
                               bricks ? bricks + (size_t)bin * vol : nullptr, bshift);
}

// ---------------------------------------------------------------------------
// FP32 spread on the matrix cores.  Per bin the brick is a sum of outer
// products,
//   brick[(x, y), z] = sum_k A_k(x, y) B_k(z),
//   A_k = round(q_k sA wx_k(x) wy_k(y)),  B_k = round(wz_k(z) kMmaSpan),
// with sA = kMmaSpan / max|q|, so that |A| and B stay below kMmaSpan < 2^23,
// each written as three signed 8-bit limbs (the limbs of v are the bytes of
// (v + 0x808080) ^ 0x808080, v = l0 + 2^8 l1 + 2^16 l2).
// v_mfma_i32_32x32x32_i8 sums the limb products over 32 particles per step:
// rows are 32 (x, y) points of the brick plane, columns the z planes.  The
// limb classes of weight 2^8 and above (li + lj >= 1) are summed exactly in
// int32; l0 x l0 (below 2^14 per particle, against products up to 2^46) is
// left out.  Every class sum is an exact integer sum, so the brick depends
// only on the set of particles whose home bin this is, not on their order:
// the bin's overflow particles are taken from the overflow list and summed
// into the same brick, and the int64 total is rounded once to the bin's
// int32 scale 2^(e - shift), the scale and shift of the LDS-atomic path.
// Operand layouts (measured): A lane l -> row l % 32, B lane l -> column
// l % 32, and byte j of lane half l / 32 is the same k slot in A and B; the
// 16 int32 results of lane l are rows 8 (r / 4) + 4 (l / 32) + r % 4, column
// l % 32.
// ---------------------------------------------------------------------------
typedef int mma_v4i __attribute__((ext_vector_type(4)));
typedef int mma_v16i __attribute__((ext_vector_type(16)));

// Brick shape of the matrix-core spread: 2n+1 = 13 points per axis in bins
// of 8 cells, so 20^3 bricks: 400 plane points in 13 row tiles of 32, 20 z
// planes in one column tile of 32.
constexpr int kMmaNZC = 13, kMmaLX = 20, kMmaLXY = 400, kMmaLZ = 20;
constexpr int kMmaTiles = (kMmaLXY + 31) / 32;
constexpr int kMmaThreads = 1024;     // 16 waves; waves below kMmaTiles own one row tile
constexpr int kMmaK = 32;             // particles per MFMA step
constexpr int kMmaMaxCall = 32768;    // int32 class sums stay exact below this
constexpr int kMmaOvfWindow = kMmaThreads;
constexpr float kMmaSpan = 8355000.0f;  // |A|, B below this: v + 0x808080 < 2^24
constexpr int kMmaNW = 3 * kMmaNZC;

// LDS layout: A table [kMmaK][kMmaLXY] packed limbs (the finished brick
// reuses it), B limb bytes [3][32][kMmaK], factors [kMmaK][3 NZC], staged
// particles, cube origins, scaled charges, overflow window and its count.
constexpr int kMmaOA = 0;
constexpr int kMmaOB = kMmaOA + kMmaK * kMmaLXY * 4;
constexpr int kMmaOW = kMmaOB + 3 * 32 * kMmaK;
constexpr int kMmaOP = (kMmaOW + kMmaK * kMmaNW * 4 + 15) & ~15;
constexpr int kMmaOO = kMmaOP + kMmaK * 16;
constexpr int kMmaOQ = kMmaOO + kMmaK * 4;
constexpr int kMmaOL = kMmaOQ + kMmaK * 4;
constexpr int kMmaON = kMmaOL + kMmaOvfWindow * 4;
constexpr int kMmaLds = kMmaON + 16;
static_assert(kMmaLXY * kMmaLZ <= kMmaK * kMmaLXY, "the brick fits the A table");
static_assert(kMmaLds <= 64 * 1024, "LDS budget");

// bytes b of four packed-limb dwords -> one operand dword (byte t from d[t])
__device__ __forceinline__ void mma_limbs4(unsigned d0, unsigned d1, unsigned d2, unsigned d3,
                                           int& l0, int& l1, int& l2) {
  const unsigned e01 = __builtin_amdgcn_perm(d1, d0, 0x05010400u);  // d0.b0 d1.b0 d0.b1 d1.b1
  const unsigned e23 = __builtin_amdgcn_perm(d3, d2, 0x05010400u);
  const unsigned f01 = __builtin_amdgcn_perm(d1, d0, 0x0c0c0602u);  // d0.b2 d1.b2 0 0
  const unsigned f23 = __builtin_amdgcn_perm(d3, d2, 0x0c0c0602u);
  l0 = (int)__builtin_amdgcn_perm(e23, e01, 0x05040100u);
  l1 = (int)__builtin_amdgcn_perm(e23, e01, 0x07060302u);
  l2 = (int)__builtin_amdgcn_perm(f23, f01, 0x05040100u);
}

__device__ __forceinline__ unsigned mma_pack(int v) {
  return ((unsigned)v + 0x808080u) ^ 0x808080u;
}

// One chunk of kc <= kMmaK staged particles (positions, cube origins and
// scaled charges in LDS): factors, the A table and B operand, then the MFMA
// steps of this wave's row tile.  Starts and ends with a barrier.
__device__ __forceinline__ void mma_chunk(const SpreadArgs& a, int kc, unsigned char* lds,
                                          mma_v16i (&acc)[4]) {
  unsigned* sA = (unsigned*)(lds + kMmaOA);
  signed char* sB = (signed char*)(lds + kMmaOB);
  float* sW = (float*)(lds + kMmaOW);
  const V4<float>* sP = (const V4<float>*)(lds + kMmaOP);
  const int* sO = (const int*)(lds + kMmaOO);
  const float* sQ = (const float*)(lds + kMmaOQ);
  const int tid = threadIdx.x;
  // the spacings as values first: a per-lane choice between fields of the
  // argument struct would become an indexed copy of the struct in scratch
  const double hx = a.g.hx, hy = a.g.hy, hz = a.g.hz;
  const double ihx = a.g.ihx, ihy = a.g.ihy, ihz = a.g.ihz;
  // factors, one thread per (particle, factor): t = (c h - x) + j h in FP64
  // (CpuGaussCharge.cpp:314-329), the exponential in FP32
  for (int e = tid; e < kc * kMmaNW; e += kMmaThreads) {
    const int k = e / kMmaNW, f = e - k * kMmaNW;
    const int axis = f / kMmaNZC, j = f - axis * kMmaNZC;
    const V4<float> p = sP[k];
    const double xv = axis == 0 ? (double)p.x : axis == 1 ? (double)p.y : (double)p.z;
    const double h = axis == 0 ? hx : axis == 1 ? hy : hz;
    const double ih = axis == 0 ? ihx : axis == 1 ? ihy : ihz;
    const int c = (int)round(xv * ih) - (kMmaNZC - 1) / 2;
    const double t = add_rn(sub_rn(mul_rn((double)c, h), xv), mul_rn((double)j, h));
    const float tc = (float)t;
    sW[e] = gexp(mul_rn(mul_rn(-tc, tc), (float)a.inv2s2));
  }
  // clear the A rows of the chunk's particles (16-byte stores)
  {
    uint4* a4 = (uint4*)sA;
    for (int i = tid; i < kc * (kMmaLXY / 4); i += kMmaThreads) a4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  // A: one thread per (particle, y offset, half of the x offsets)
  for (int e = tid; e < kc * kMmaNZC * 2; e += kMmaThreads) {
    const int kr = e >> 1, hx = e & 1;
    const int k = kr / kMmaNZC, dy = kr - k * kMmaNZC;
    const int o = sO[k];
    const float* w = sW + k * kMmaNW;
    const float qy = sQ[k] * w[kMmaNZC + dy];
    unsigned* row = sA + k * kMmaLXY + (o & 255) + kMmaLX * (((o >> 8) & 255) + dy);
    const int dx0 = hx ? 7 : 0, dx1 = hx ? kMmaNZC : 7;
#pragma unroll
    for (int dx = 0; dx < 7; dx++)
      if (dx0 + dx < dx1) row[dx0 + dx] = mma_pack((int)rintf(qy * w[dx0 + dx]));
  }
  // B: limb bytes [limb][z][slot]; zero outside the particle's z window and
  // for slots past kc
  for (int e = tid; e < 32 * kMmaK; e += kMmaThreads) {
    const int z = e / kMmaK, k = e - z * kMmaK;
    int v = 0;
    if (k < kc && z < kMmaLZ) {
      const int dz = z - (sO[k] >> 16);
      if (dz >= 0 && dz < kMmaNZC) v = (int)rintf(sW[k * kMmaNW + 2 * kMmaNZC + dz] * kMmaSpan);
    }
    const unsigned pk = mma_pack(v);
    sB[(0 * 32 + z) * kMmaK + k] = (signed char)(pk & 255u);
    sB[(1 * 32 + z) * kMmaK + k] = (signed char)((pk >> 8) & 255u);
    sB[(2 * 32 + z) * kMmaK + k] = (signed char)((pk >> 16) & 255u);
  }
  __syncthreads();
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (wid < kMmaTiles) {
    const int lane = tid & 63, half = lane >> 5, col = lane & 31;
    const mma_v4i b0 = *(const mma_v4i*)(sB + (0 * 32 + col) * kMmaK + 16 * half);
    const mma_v4i b1 = *(const mma_v4i*)(sB + (1 * 32 + col) * kMmaK + 16 * half);
    const mma_v4i b2 = *(const mma_v4i*)(sB + (2 * 32 + col) * kMmaK + 16 * half);
    const int m = 32 * wid + col;
    const unsigned* ak = sA + 16 * half * kMmaLXY + m;
    unsigned d[16];
#pragma unroll
    for (int j = 0; j < 16; j++) d[j] = m < kMmaLXY ? ak[j * kMmaLXY] : 0u;
    mma_v4i a0, a1, a2;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int l0, l1, l2;
      mma_limbs4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3], l0, l1, l2);
      a0[q] = l0;
      a1[q] = l1;
      a2[q] = l2;
    }
    // limb classes of weight 2^8 (a0 b1 + a1 b0), 2^16 (a0 b2 + a1 b1 + a2 b0),
    // 2^24 (a1 b2 + a2 b1) and 2^32 (a2 b2)
    acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b1, acc[0], 0, 0, 0);
    acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b2, acc[1], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, acc[1], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, b0, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b2, acc[2], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, b1, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, b2, acc[3], 0, 0, 0);
  }
  __syncthreads();
}

// Cube origin of a particle relative to its home bin (as k_bin stores it).
__device__ __forceinline__ int mma_origin(const Geom& g, const BinShape& b, const V4<float>& p) {
  const int ox = wrapi((int)round((double)p.x * g.ihx) - g.nx, g.X);
  const int oy = wrapi((int)round((double)p.y * g.ihy) - g.ny, g.Y);
  const int oz = wrapi((int)round((double)p.z * g.ihz) - g.nz, g.Z);
  return (ox % b.Bx) | ((oy % b.By) << 8) | ((oz % b.Bz) << 16);
}

// A bin with more than kMmaMaxCall particles: every one of them into the
// int64 grid at the bin's scale, one wave per particle
__device__ __forceinline__ void mma_giant_bin(
    const SpreadArgs& a, const V4<float>* __restrict__ xyzq,
    const V4<float>* __restrict__ sorted, const int* __restrict__ ovf, int novf, int bin,
    int cnt, int eg, int shift, unsigned long long* __restrict__ grid64) {
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int nwv = kMmaThreads / 64;
  const double S = ldexp(1.0, eg - shift);
  for (int i = wid; i < cnt; i += nwv)
    spread_one_global<float>(a.g, a.inv2s2, a.pref, S, 0, a.pref1,
                             sorted[(size_t)bin * a.cap_bin + i], grid64, shift);
  for (int i = wid; i < novf; i += nwv) {
    const V4<float> p = xyzq[ovf[i]];
    if (home_bin<float>(a.g, a.b, p) == bin)
      spread_one_global<float>(a.g, a.inv2s2, a.pref, S, 0, a.pref1, p, grid64, shift);
  }
}

// One workgroup per bin (no extra overflow workgroups: each bin sums its own
// overflow particles).  A bin holding more than kMmaMaxCall particles takes
// the per-contribution path instead: its particles go into the int64 grid at
// its scale, as overflow particles of the LDS-atomic path do, and its brick
// is marked empty.
__global__ void __launch_bounds__(kMmaThreads)
    k_spread_mma(SpreadArgs a, const V4<float>* __restrict__ xyzq,
                 const V4<float>* __restrict__ sorted, const int* __restrict__ counts,
                 int* __restrict__ counts_other, const BinStats* __restrict__ stats,
                 BinStats* __restrict__ stats_other, const int* __restrict__ ovf,
                 unsigned long long* __restrict__ grid64, unsigned* __restrict__ bricks,
                 int* __restrict__ bshift, const int* __restrict__ loc) {
  extern __shared__ __align__(16) unsigned char mma_smem[];
  const BinStats st = *stats;
  const int eg = fixed_exponent(st.qmax_bits, a.n - st.invalid, a.images, a.pref_bound);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stats_other->qmax_bits = 0ull;
    stats_other->invalid = 0;
    stats_other->overflow = 0;
    stats_other->spill = 0;
  }
  const int pos = xcd_remap(blockIdx.x, a.nwg_bins);
  if (pos >= a.b.nbins) return;
  const int bin = bin_at(a.order, pos);
  const int call = counts[bin];
  const int cnt = min(call, a.cap_bin);
  if (threadIdx.x == 0) {
    counts_other[bin] = 0;
    if (call == 0 || call > kMmaMaxCall) bshift[bin] = -1;  // no brick this step
  }
  if (call == 0) return;
  const Geom& g = a.g;
  const BinShape& b = a.b;
  const int ebin = fixed_exponent_bin(st.qmax_bits, call, a.pref_bound);
  const int shift = eg > ebin ? eg - ebin : 0;
  const int tid = threadIdx.x;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (call > kMmaMaxCall) {  // per-contribution path, one wave per particle
    mma_giant_bin(a, xyzq, sorted, ovf, st.overflow, bin, cnt, eg, shift, grid64);
    return;
  }
  // A = q qscale wx wy with qscale = kMmaSpan / qmax, so |A| <= kMmaSpan; the
  // total T = sum A B is q wx wy wz at the scale qscale kMmaSpan and goes to
  // the brick's q pref wx wy wz 2^(eg - shift), rounded once
  union {
    unsigned long long u;
    double d;
  } qv;
  qv.u = st.qmax_bits;
  const float qscale = qv.d > 0.0 ? (float)((double)kMmaSpan / qv.d) : 0.0f;
  const double fct = qv.d > 0.0 ? ldexp(1.0, eg - shift) * a.pref /
                                      ((double)qscale * (double)kMmaSpan)
                                : 0.0;

  unsigned char* lds = mma_smem;
  V4<float>* sP = (V4<float>*)(lds + kMmaOP);
  int* sO = (int*)(lds + kMmaOO);
  float* sQ = (float*)(lds + kMmaOQ);
  int* sL = (int*)(lds + kMmaOL);
  int* sN = (int*)(lds + kMmaON);

  mma_v16i acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[c][r] = 0;

  // the bin's slots
  const V4<float>* bs = sorted + (size_t)bin * a.cap_bin;
  const int* bl = loc + (size_t)bin * a.cap_bin;
  for (int c0 = 0; c0 < cnt; c0 += kMmaK) {
    const int kc = min(kMmaK, cnt - c0);
    if (tid < kc) {
      const V4<float> p = bs[c0 + tid];
      sP[tid] = p;
      sO[tid] = bl[c0 + tid];
      sQ[tid] = p.w * qscale;
    }
    __syncthreads();
    mma_chunk(a, kc, lds, acc);
  }
  // the bin's overflow particles, found on the overflow list a window at a time
  if (call > cnt) {
    for (int w0 = 0; w0 < st.overflow; w0 += kMmaOvfWindow) {
      if (tid == 0) *sN = 0;
      __syncthreads();
      const int i = w0 + tid;
      if (i < st.overflow) {
        const int pi = ovf[i];
        if (home_bin<float>(g, b, xyzq[pi]) == bin) sL[atomicAdd(sN, 1)] = pi;
      }
      __syncthreads();
      const int nm = *sN;
      __syncthreads();  // every thread has read the count before it is reset
      for (int c0 = 0; c0 < nm; c0 += kMmaK) {
        const int kc = min(kMmaK, nm - c0);
        if (tid < kc) {
          const V4<float> p = xyzq[sL[c0 + tid]];
          sP[tid] = p;
          sO[tid] = mma_origin(g, b, p);
          sQ[tid] = p.w * qscale;
        }
        __syncthreads();
        mma_chunk(a, kc, lds, acc);
      }
    }
  }
  // exact class sums -> the brick at 2^(eg - shift): each class sum converts
  // to FP64 exactly, the weighted sum is formed in a fixed order and rounded
  // once (the same operations for any order of the particles)
  unsigned* sOut = (unsigned*)(lds + kMmaOA);
  if (wid < kMmaTiles) {
    const int lane = tid & 63, half = lane >> 5, col = lane & 31;
    const double f1 = fct * 256.0, f2 = fct * 65536.0, f3 = fct * 16777216.0,
                 f4 = fct * 4294967296.0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int m = 32 * wid + 8 * (r >> 2) + 4 * half + (r & 3);
      if (m < kMmaLXY && col < kMmaLZ) {
        double v = (double)acc[3][r] * f4;
        v = __builtin_fma((double)acc[2][r], f3, v);
        v = __builtin_fma((double)acc[1][r], f2, v);
        v = __builtin_fma((double)acc[0][r], f1, v);
        sOut[m + kMmaLXY * col] = (unsigned)(int)rint(v);
      }
    }
  }
  __syncthreads();
  constexpr int vol = kMmaLXY * kMmaLZ;
  uint4* out = (uint4*)(bricks + (size_t)bin * vol);
  for (int j = tid; j < vol / 4; j += kMmaThreads) out[j] = ((const uint4*)sOut)[j];
  if (tid == 0) bshift[bin] = shift;
}

// The matrix-core spread applies to FP32 grids with int32 bricks of
// 20 x 20 x 20 (bins of 8 cells, 2n+1 = 13) and no normalisation.
bool spread_mma_ok(bool dbl, const Geom& g, const BinShape& b, bool normalize, bool acc32,
                   bool bricks) {
  if (dbl || !acc32 || !bricks || normalize || b.direct) return false;
  if (2 * g.nx + 1 != kMmaNZC || 2 * g.ny + 1 != kMmaNZC || 2 * g.nz + 1 != kMmaNZC) return false;
  return b.Lx == kMmaLX && b.Ly * b.Lx == kMmaLXY && b.Lz == kMmaLZ;
}

// Bricks covering grid coordinate p on one axis: every (bin b, brick cell l)
// with (b*B + l) == p mod M, l < L.  Returns the count (at most KR).
constexpr int KR = 8;
__device__ __forceinline__ int brick_axis(int p, int M, int B, int nb, int L, int (&bl)[KR],
                                          int (&ll)[KR]) {
  int c = 0;
  for (int q = p; q - (nb - 1) * B < L; q += M) {
    int blo = q - L + 1;
    blo = blo <= 0 ? 0 : (blo + B - 1) / B;
    const int bhi = min(q / B, nb - 1);
    for (int b = blo; b <= bhi && c < KR; b++) {
      bl[c] = b;
      ll[c] = q - b * B;
      c++;
    }
  }
  return c;
}

// Grid = sum of the bricks covering each point (int64, exact, fixed order)
// plus the overflow particles' int64 grid, times 2^-e; the int64 grid is
// cleared again where it was used.
template <typename T, typename ACC>
__global__ void __launch_bounds__(256)
    k_brick_reduce(Geom g, BinShape b, const ACC* __restrict__ bricks,
                   const int* __restrict__ bshift, const BinStats* __restrict__ stats, int n,
                   int images, double pref_bound, unsigned long long* __restrict__ grid64,
                   T* __restrict__ out) {
  const BinStats st = *stats;
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, n - st.invalid, images, pref_bound));
  const bool ovf = st.overflow > 0;
  const size_t vol = (size_t)b.Lx * b.Ly * b.Lz;
  const size_t np = (size_t)g.X * g.Y * g.Z;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < np;
       p += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(p % g.X);
    const size_t t = p / g.X;
    const int y = (int)(t % g.Y), z = (int)(t / g.Y);
    int bx[KR], lx[KR], by[KR], ly[KR], bz[KR], lz[KR];
    const int nx = brick_axis(x, g.X, b.Bx, b.nbx, b.Lx, bx, lx);
    const int ny = brick_axis(y, g.Y, b.By, b.nby, b.Ly, by, ly);
    const int nz = brick_axis(z, g.Z, b.Bz, b.nbz, b.Lz, bz, lz);
    long long sum = 0;
    for (int iz = 0; iz < nz; iz++)
      for (int iy = 0; iy < ny; iy++) {
        const int brow = b.nbx * (by[iy] + b.nby * bz[iz]);
        const int crow = b.Lx * (ly[iy] + b.Ly * lz[iz]);
        for (int ix = 0; ix < nx; ix++) {
          const int bin = bx[ix] + brow;
          const int sh = bshift[bin];
          if (sh < 0) continue;
          const ACC raw = bricks[(size_t)bin * vol + lx[ix] + crow];
          sum += sizeof(ACC) == 4 ? ((long long)(int)raw << sh) : (long long)raw;
        }
      }
    if (ovf) {
      sum += (long long)grid64[p];
      grid64[p] = 0ull;
    }
    out[p] = (T)((double)sum * inv);
  }
}

// Column-owner spread: thread j owns brick column (x, y) = (j % Lx, j / Lx)
// and keeps its Lz = Bz + 2n values in registers, as int32 (FP32 grids, per-bin
// scale) or int64 (global scale).  All threads walk all staged particles of
// the bin; a thread inside a particle's (x, y) footprint adds
// round(qxy * wz[t] * 2^e) for the particle's 2n+1 z planes.  fma(a, b, M)
// with M = 1.5*2^52 rounds the exact product to an integer held in the low
// mantissa word, so one FMA and one integer add make one update; the
// particle's z origin is wave-uniform, so the switch keeps every register
// index static.  No LDS atomics; the column is then added into the global
// int64 grid.
constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr long long kMagicBits = 0x4338000000000000LL;

template <typename ACC, int NZC, int BZC>
struct ColAcc {
  ACC v[BZC + NZC - 1];
};

template <typename ACC, int NZC, int OZ, int BZC>
__device__ __forceinline__ void col_add(ColAcc<ACC, NZC, BZC>& c, double qxy, const double* wz) {
#pragma unroll
  for (int t = 0; t < NZC; t++) {
    const double d = __builtin_fma(qxy, wz[t], kMagic);
    if (sizeof(ACC) == 4)
      c.v[OZ + t] += (ACC)__double2loint(d);
    else
      c.v[OZ + t] += (ACC)(__double_as_longlong(d) - kMagicBits);
  }
}

template <typename CT, int NZC, int BZC, int BS, typename ACC>
__global__ void __launch_bounds__(BS)
    k_spread_cols(SpreadArgs a, const V4<CT>* __restrict__ xyzq, const V4<CT>* __restrict__ sorted,
                  const int* __restrict__ counts, int* __restrict__ counts_other,
                  const BinStats* __restrict__ stats, BinStats* __restrict__ stats_other,
                  const int* __restrict__ ovf, unsigned long long* __restrict__ grid64) {
  extern __shared__ __align__(32) unsigned char cols_smem[];
  const Geom& g = a.g;
  const BinShape& b = a.b;
  const BinStats st = *stats;
  const int eg = fixed_exponent(st.qmax_bits, a.n - st.invalid, a.images, a.pref_bound);
  const double S = ldexp(1.0, eg);
  if ((int)blockIdx.x >= a.nwg_bins) {  // overflow particles, global atomics
    if (blockIdx.x == (unsigned)a.nwg_bins && threadIdx.x == 0) {
      stats_other->qmax_bits = 0ull;
      stats_other->invalid = 0;
      stats_other->overflow = 0;
      stats_other->spill = 0;
    }
    const int wg = blockIdx.x - a.nwg_bins;
    const int gw = wg * (BS / 64) + (threadIdx.x >> 6);
    for (int k = __builtin_amdgcn_readfirstlane(gw); k < st.overflow; k += kOverflowWG * (BS / 64))
      spread_overflow<CT, ACC>(a, eg, st.qmax_bits, counts, xyzq[ovf[k]], grid64);
    return;
  }
  const int pos = xcd_remap(blockIdx.x, a.nwg_bins);
  if (pos >= b.nbins) return;
  const int bin = bin_at(a.order, pos);
  const int call = counts[bin];
  const int cnt = min(call, a.cap_bin);
  if (threadIdx.x == 0) counts_other[bin] = 0;
  if (cnt == 0) return;
  int shift = 0;
  double Sb = S;
  if (sizeof(ACC) == 4) {  // int32 columns: per-bin exponent keeps each sum below 2^30
    const int eb = fixed_exponent_bin(st.qmax_bits, call, a.pref_bound);
    shift = eg > eb ? eg - eb : 0;
    Sb = ldexp(S, -shift);
  }
  const int s0 = bin * a.cap_bin, s1 = s0 + cnt;
  const int bxi = bin % b.nbx, byi = (bin / b.nbx) % b.nby, bzi = bin / (b.nbx * b.nby);
  const int Lx = b.Lx, Ly = b.Ly;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1;
  constexpr int NW_MAX = 64;
  const int NW = NX + NY + NZC;
  const int cap = a.cap;
  const SpreadLds lay = spread_lds(0, cap, NW, sizeof(CT));
  V4<CT>* sX = (V4<CT>*)(cols_smem + lay.oX);
  CT* sW = (CT*)(cols_smem + lay.oW);
  double* sQ = (double*)(cols_smem + lay.oQ);
  int* sO = (int*)(cols_smem + lay.oB);  // packed local origin ox | oy<<10 | oz<<20
  (void)NW_MAX;

  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int nwv = BS / 64;
  LaneRole role;
  if (lane < NX) {
    role.axis = 0; role.j = lane;
  } else if (lane < NX + NY) {
    role.axis = 1; role.j = lane - NX;
  } else if (lane < NW) {
    role.axis = 2; role.j = lane - NX - NY;
  } else {
    role.axis = -1; role.j = 0;
  }
  const int col = threadIdx.x;
  const bool has_col = col < Lx * Ly;
  const int cyl = has_col ? col / Lx : 0;
  const int cxl = has_col ? col - cyl * Lx : 0;
  ColAcc<ACC, NZC, BZC> acc;
#pragma unroll
  for (int z = 0; z < BZC + NZC - 1; z++) acc.v[z] = 0;

  for (int c0 = s0; c0 < s1; c0 += cap) {
    const int na = min(cap, s1 - c0);
    for (int t = threadIdx.x; t < na; t += BS) sX[t] = sorted[c0 + t];
    __syncthreads();
    for (int ai = wid; ai < na; ai += nwv) {
      const V4<CT> p = sX[ai];
      const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
      const int cx = (int)round(x * g.ihx) - g.nx;
      const int cy = (int)round(y * g.ihy) - g.ny;
      const int cz = (int)round(z * g.ihz) - g.nz;
      if (role.axis >= 0)
        sW[(size_t)ai * NW + lane] = lane_weight<CT>(role, cx, cy, cz, x, y, z, g, a.inv2s2);
      if (lane == 0) {
        const int ox = wrapi(cx, g.X) - bxi * b.Bx;  // in [0, Bx)
        const int oy = wrapi(cy, g.Y) - byi * b.By;
        const int oz = wrapi(cz, g.Z) - bzi * b.Bz;  // in [0, Bz) <= BZC
        sO[ai] = ox | (oy << 10) | (oz << 20);
        sQ[ai] = ((double)p.w * a.pref) * Sb;
      }
    }
    __syncthreads();
    if (a.normalize) {  // CpuGaussCharge.cpp:330-362: per-axis sums in j order, in CT
      for (int e = threadIdx.x; e < na * 3; e += BS) {
        const int ai = e / 3, ax = e - 3 * ai;
        const int j0 = ax == 0 ? 0 : ax == 1 ? NX : NX + NY;
        const int nj = ax == 0 ? NX : ax == 1 ? NY : NZC;
        CT* w = sW + (size_t)ai * NW + j0;
        CT sum = 0;
        for (int j = 0; j < nj; j++) sum += w[j];
        const CT nrm = (CT)a.pref1 / sum;
        for (int j = 0; j < nj; j++) w[j] = w[j] * nrm;
      }
      __syncthreads();
    }
    for (int ai = 0; ai < na; ai++) {
      const int o = sO[ai];
      const int tx = cxl - (o & 1023);
      const int ty = cyl - ((o >> 10) & 1023);
      if (has_col && (unsigned)tx < (unsigned)NX && (unsigned)ty < (unsigned)NY) {
        const CT* w = sW + (size_t)ai * NW;
        // ((q pref 2^e) wx) wy: the reference's product order, scaled exactly
        const double qxy = (sQ[ai] * (double)w[tx]) * (double)w[NX + ty];
        double wz[NZC];
#pragma unroll
        for (int t = 0; t < NZC; t++) wz[t] = (double)w[NX + NY + t];
        switch (o >> 20) {
#define GSE_COL_CASE(OZ) \
  case OZ:               \
    if (OZ < BZC) col_add<ACC, NZC, (OZ < BZC ? OZ : 0), BZC>(acc, qxy, wz); \
    break;
          GSE_COL_CASE(0)
          GSE_COL_CASE(1)
          GSE_COL_CASE(2)
          GSE_COL_CASE(3)
          GSE_COL_CASE(4)
          GSE_COL_CASE(5)
          GSE_COL_CASE(6)
          GSE_COL_CASE(7)
#undef GSE_COL_CASE
          default:
            break;
        }
      }
    }
    __syncthreads();
  }

  if (has_col) {
    const int X = g.X, Y = g.Y, Z = g.Z;
    int gx = bxi * b.Bx + cxl, gy = byi * b.By + cyl;
    while (gx >= X) gx -= X;
    while (gy >= Y) gy -= Y;
    int gz = bzi * b.Bz;
    unsigned long long* colp = grid64 + gx + (size_t)X * gy;
    const size_t plane = (size_t)X * Y;
    const int Lz = b.Bz + NZC - 1;
#pragma unroll
    for (int z = 0; z < BZC + NZC - 1; z++) {
      if (z < Lz) {
        const unsigned long long v =
            sizeof(ACC) == 4 ? (unsigned long long)((long long)(int)acc.v[z] << shift)
                             : (unsigned long long)acc.v[z];
        int zz = gz + z;
        while (zz >= Z) zz -= Z;
        if (v != 0ull) atomicAdd(&colp[(size_t)zz * plane], v);
      }
    }
  }
}

// Supports too wide for the LDS path: all binned particles (one bin of
// capacity n) through the global path, one wave per particle.
template <typename CT>
__global__ void __launch_bounds__(256)
    k_spread_direct(SpreadArgs a, const V4<CT>* __restrict__ sorted, const int* __restrict__ counts,
                    int* __restrict__ counts_other, const BinStats* __restrict__ stats,
                    BinStats* __restrict__ stats_other, unsigned long long* __restrict__ grid64) {
  const BinStats st = *stats;
  const double S = ldexp(1.0, fixed_exponent(st.qmax_bits, a.n - st.invalid, a.images,
                                             a.pref_bound));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts_other[0] = 0;
    stats_other->qmax_bits = 0ull;
    stats_other->invalid = 0;
    stats_other->overflow = 0;
    stats_other->spill = 0;
  }
  const int n = min(counts[0], a.cap_bin);
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (gridDim.x * blockDim.x) >> 6;
  for (int i = gw; i < n; i += nwave)
    spread_one_global<CT>(a.g, a.inv2s2, a.pref, S, a.normalize, a.pref1, sorted[i], grid64);
}

template <typename T>
__global__ void k_fixed_to_grid(unsigned long long* __restrict__ g64, T* __restrict__ out,
                                size_t count, const BinStats* __restrict__ stats, int n,
                                int images, double pref_bound, int clear) {
  const BinStats st = *stats;
  const double inv = ldexp(1.0, -fixed_exponent(st.qmax_bits, n - st.invalid, images, pref_bound));
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x) {
    const long long v = (long long)g64[i];
    out[i] = (T)((double)v * inv);
    if (clear) g64[i] = 0ull;
  }
}

__global__ void k_fill_u64(unsigned long long* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

template <typename T>
__global__ void k_fill(T* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

static unsigned grid_for(size_t count, unsigned threads, unsigned cap = 4096) {
  size_t g = (count + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Tiled form of k_brick_reduce for bins of 8^3 cells on grids that are whole
// numbers of bins (>= 3 per axis) with bricks at most 3 bins wide: one
// 512-thread workgroup per bin interior, each thread one grid point; the
// bricks of the bin and of its neighbours at -1, -2 bins that cover it are
// walked as a uniform loop, so brick shifts are scalar loads and each brick
// row is read as 8 consecutive values.
// One bin interior of one system (bricks, bshift, grid64 and out at the
// system's base); inv = 2^-e, ovf: the overflow int64 grid holds data.
template <typename T, typename ACC>
__device__ __forceinline__ void reduce_tile_bin(const Geom& g, const BinShape& b, int bin,
                                                const ACC* __restrict__ bricks,
                                                const int* __restrict__ bshift, double inv,
                                                bool ovf, unsigned long long* __restrict__ grid64,
                                                T* __restrict__ out) {
  const int cx = bin % b.nbx, cy = (bin / b.nbx) % b.nby, cz = bin / (b.nbx * b.nby);
  const int tx = threadIdx.x & 7, ty = (threadIdx.x >> 3) & 7, tz = threadIdx.x >> 6;
  const int Lx = b.Lx, Ly = b.Ly, Lz = b.Lz;
  const size_t vol = (size_t)Lx * Ly * Lz;
  // the 27 candidate bricks (the bin and its neighbours at -1, -2 bins on
  // each axis), one z layer of 9 at a time: a layer's shifts and values are
  // all loaded before any is summed, so its loads overlap; each load fills a
  // zeroed value under `if (ok)`, which keeps the register count low
  long long sum = 0;
#pragma unroll
  for (int dz = 0; dz < 3; dz++) {
    int sh[9];
    ACC raw[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int dx = k % 3, dy = k / 3;
      int bx = cx - dx, by = cy - dy, bz = cz - dz;
      bx += bx < 0 ? b.nbx : 0;
      by += by < 0 ? b.nby : 0;
      bz += bz < 0 ? b.nbz : 0;
      const int nbin = bx + b.nbx * (by + b.nby * bz);
      sh[k] = bshift[nbin];
      const int lx = tx + 8 * dx, ly = ty + 8 * dy, lz = tz + 8 * dz;
      const bool ok = sh[k] >= 0 && lx < Lx && ly < Ly && lz < Lz;
      ACC v = 0;
      if (ok) v = bricks[(size_t)nbin * vol + lx + Lx * (ly + Ly * lz)];
      raw[k] = v;
    }
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (sh[k] >= 0)
        sum += sizeof(ACC) == 4 ? ((long long)(int)raw[k] << sh[k]) : (long long)raw[k];
  }
  const size_t p = (size_t)(8 * cx + tx) + (size_t)g.X * ((8 * cy + ty) + (size_t)g.Y * (8 * cz + tz));
  if (ovf) {
    sum += (long long)grid64[p];
    grid64[p] = 0ull;
  }
  out[p] = (T)((double)sum * inv);
}

template <typename T, typename ACC>
__global__ void __launch_bounds__(512)
    k_brick_reduce_tile(Geom g, BinShape b, int nwg, const ACC* __restrict__ bricks,
                        const int* __restrict__ bshift, const BinStats* __restrict__ stats, int n,
                        int images, double pref_bound, unsigned long long* __restrict__ grid64,
                        T* __restrict__ out) {
  const int bin = xcd_remap(blockIdx.x, nwg);
  if (bin >= b.nbins) return;
  const BinStats st = *stats;
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, n - st.invalid, images, pref_bound));
  reduce_tile_bin<T, ACC>(g, b, bin, bricks, bshift, inv, st.overflow > 0, grid64, out);
}

// k_brick_reduce_tile with 32-bit bricks whose rows are whole 16-byte
// pieces (Lx a multiple of 4): 128 threads per bin interior, each thread four
// consecutive x points read and written as 16-byte pieces, the 27 candidate
// bricks taken one z layer (9 bricks) at a time.  Each load fills a
// zero-initialised vector under `if (ok)`; selecting between a load and a
// zero vector instead took 94 VGPRs (5 waves per SIMD) against 54 (8).
// Every brick word is read exactly once; NT makes the brick loads
// non-temporal, which measured slower at every grid size from 80^3 to 216^3
// (a brick row's three pieces go to three tiles, and the later pieces then
// miss), so it is off by default.
typedef unsigned BrickWord4 __attribute__((ext_vector_type(4)));
template <typename T, bool NT>
__device__ __forceinline__ void reduce_tile4_bin(const Geom& g, const BinShape& b, int bin,
                                                 const unsigned* __restrict__ bricks,
                                                 const int* __restrict__ bshift, double inv,
                                                 bool ovf,
                                                 unsigned long long* __restrict__ grid64,
                                                 T* __restrict__ out) {
  const int cx = bin % b.nbx, cy = (bin / b.nbx) % b.nby, cz = bin / (b.nbx * b.nby);
  const int tq = threadIdx.x & 1, ty = (threadIdx.x >> 1) & 7, tz = threadIdx.x >> 4;
  const int Lx = b.Lx, Ly = b.Ly, Lz = b.Lz;
  const size_t vol = (size_t)Lx * Ly * Lz;
  long long sum[4] = {0, 0, 0, 0};
#pragma unroll
  for (int dz = 0; dz < 3; dz++) {
    int sh[9];
    uint4 raw[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int dx = k % 3, dy = k / 3;
      int bx = cx - dx, by = cy - dy, bz = cz - dz;
      bx += bx < 0 ? b.nbx : 0;
      by += by < 0 ? b.nby : 0;
      bz += bz < 0 ? b.nbz : 0;
      const int nbin = bx + b.nbx * (by + b.nby * bz);
      sh[k] = bshift[nbin];
      const int lx = 4 * tq + 8 * dx, ly = ty + 8 * dy, lz = tz + 8 * dz;
      const bool ok = sh[k] >= 0 && lx < Lx && ly < Ly && lz < Lz;
      const BrickWord4* src =
          (const BrickWord4*)&bricks[(size_t)nbin * vol + lx + Lx * (ly + Ly * lz)];
      BrickWord4 r4 = {0u, 0u, 0u, 0u};
      if (ok) r4 = NT ? __builtin_nontemporal_load(src) : *src;
      raw[k] = make_uint4(r4.x, r4.y, r4.z, r4.w);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
      if (sh[k] < 0) continue;
      sum[0] += (long long)(int)raw[k].x << sh[k];
      sum[1] += (long long)(int)raw[k].y << sh[k];
      sum[2] += (long long)(int)raw[k].z << sh[k];
      sum[3] += (long long)(int)raw[k].w << sh[k];
    }
  }
  const size_t p = (size_t)(8 * cx + 4 * tq) + (size_t)g.X * ((8 * cy + ty) + (size_t)g.Y * (8 * cz + tz));
  if (ovf) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      sum[c] += (long long)grid64[p + c];
      grid64[p + c] = 0ull;
    }
  }
  T o4[4];
#pragma unroll
  for (int c = 0; c < 4; c++) o4[c] = (T)((double)sum[c] * inv);
  if (sizeof(T) == 4) {
    *(float4*)&out[p] = make_float4((float)o4[0], (float)o4[1], (float)o4[2], (float)o4[3]);
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++) out[p + c] = o4[c];
  }
}

template <typename T, bool NT>
__global__ void __launch_bounds__(128)
    k_brick_reduce_tile4(Geom g, BinShape b, int nwg, const unsigned* __restrict__ bricks,
                         const int* __restrict__ bshift, const BinStats* __restrict__ stats,
                         int n, int images, double pref_bound,
                         unsigned long long* __restrict__ grid64, T* __restrict__ out) {
  const int bin = xcd_remap(blockIdx.x, nwg);
  if (bin >= b.nbins) return;
  const BinStats st = *stats;
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, n - st.invalid, images, pref_bound));
  reduce_tile4_bin<T, NT>(g, b, bin, bricks, bshift, inv, st.overflow > 0, grid64, out);
}

// k_brick_reduce_tile with 64-bit bricks whose rows have an even length: 256
// threads per bin interior, each thread two consecutive x points read as one
// 16-byte piece per brick, the 27 candidate bricks taken one z layer (9
// bricks) at a time as in reduce_tile4_bin.
template <typename T>
__device__ __forceinline__ void reduce_tile2_bin(const Geom& g, const BinShape& b, int bin,
                                                 const unsigned long long* __restrict__ bricks,
                                                 const int* __restrict__ bshift, double inv,
                                                 bool ovf,
                                                 unsigned long long* __restrict__ grid64,
                                                 T* __restrict__ out) {
  typedef unsigned long long BrickWord2 __attribute__((ext_vector_type(2)));
  const int cx = bin % b.nbx, cy = (bin / b.nbx) % b.nby, cz = bin / (b.nbx * b.nby);
  const int tq = threadIdx.x & 3, ty = (threadIdx.x >> 2) & 7, tz = threadIdx.x >> 5;
  const int Lx = b.Lx, Ly = b.Ly, Lz = b.Lz;
  const size_t vol = (size_t)Lx * Ly * Lz;
  long long sum[2] = {0, 0};
#pragma unroll
  for (int dz = 0; dz < 3; dz++) {
    BrickWord2 raw[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int dx = k % 3, dy = k / 3;
      int bx = cx - dx, by = cy - dy, bz = cz - dz;
      bx += bx < 0 ? b.nbx : 0;
      by += by < 0 ? b.nby : 0;
      bz += bz < 0 ? b.nbz : 0;
      const int nbin = bx + b.nbx * (by + b.nby * bz);
      const int lx = 2 * tq + 8 * dx, ly = ty + 8 * dy, lz = tz + 8 * dz;
      const bool ok = bshift[nbin] >= 0 && lx < Lx && ly < Ly && lz < Lz;
      BrickWord2 v = {0ull, 0ull};
      if (ok) v = *(const BrickWord2*)&bricks[(size_t)nbin * vol + lx + Lx * (ly + Ly * lz)];
      raw[k] = v;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
      sum[0] += (long long)raw[k].x;
      sum[1] += (long long)raw[k].y;
    }
  }
  const size_t p = (size_t)(8 * cx + 2 * tq) + (size_t)g.X * ((8 * cy + ty) + (size_t)g.Y * (8 * cz + tz));
  if (ovf) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      sum[c] += (long long)grid64[p + c];
      grid64[p + c] = 0ull;
    }
  }
  out[p] = (T)((double)sum[0] * inv);
  out[p + 1] = (T)((double)sum[1] * inv);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_brick_reduce_tile2(Geom g, BinShape b, int nwg, const unsigned long long* __restrict__ bricks,
                         const int* __restrict__ bshift, const BinStats* __restrict__ stats,
                         int n, int images, double pref_bound,
                         unsigned long long* __restrict__ grid64, T* __restrict__ out) {
  const int bin = xcd_remap(blockIdx.x, nwg);
  if (bin >= b.nbins) return;
  const BinStats st = *stats;
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, n - st.invalid, images, pref_bound));
  reduce_tile2_bin<T>(g, b, bin, bricks, bshift, inv, st.overflow > 0, grid64, out);
}

// ---------------------------------------------------------------------------
// Batched independent systems: the same per-bin bodies, one launch for all
// systems.  Spread and gather walk a table of (system, bin) items with a
// fixed grid of workgroups (each XCD takes one contiguous range of items, so
// the bins of one system meet in one L2).
// ---------------------------------------------------------------------------
// System owning concatenated particle i: largest s with pofs[s] <= i.
__device__ __forceinline__ int batch_system(const int* __restrict__ pofs, int nsys, int i) {
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pofs[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Item it of the XCD-aware persistent walk: workgroup w runs on XCD w % 8
// and takes the items of that XCD's contiguous range in steps of the number
// of workgroups per XCD.
struct ItemWalk {
  int it, end, step;
};
__device__ __forceinline__ ItemWalk item_walk(int nitems) {
  const int nx = (int)(gridDim.x >> 3);  // workgroups per XCD (grid is a multiple of 8)
  const int x = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3);
  const int chunk = (nitems + 7) >> 3;
  const int b = x * chunk;
  const int e = min(nitems, b + chunk);
  return ItemWalk{b + j, e, nx};
}

__device__ __forceinline__ SpreadArgs batch_spread_args(const BatchSys& S) {
  SpreadArgs a;
  a.g = S.g;
  a.b = S.b;
  a.inv2s2 = S.inv2s2;
  a.pref = S.pref;
  a.pref1 = 0.0;
  a.normalize = 0;
  a.cap = S.spread_cap;
  a.cap_bin = S.cap_bin;
  a.n = S.n;
  a.images = S.images;
  a.pref_bound = S.pref_bound;
  a.nwg_bins = 0;
  a.order = nullptr;
  return a;
}

__global__ void k_batch_clear(int* counts, int nbins, BinStats* stats, int nsys, int* novf) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (int i = i0; i < nbins; i += st) counts[i] = 0;
  for (int i = i0; i < nsys; i += st) {
    stats[i].qmax_bits = 0ull;
    stats[i].invalid = 0;
    stats[i].overflow = 0;
    stats[i].spill = 0;
  }
  if (i0 == 0) *novf = 0;
}

template <typename CT>
__global__ void __launch_bounds__(256)
    k_batch_bin(BatchDev d, const V4<CT>* __restrict__ xyzq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int s = -1;
  double aq = 0.0;
  if (i < d.ntot) {
    s = batch_system(d.pofs, d.nsys, i);
    const BatchSys& S = d.sys[s];
    const Geom& g = S.g;
    const V4<CT> p = xyzq[i];
    const double fx = (double)p.x * g.ihx;
    const double fy = (double)p.y * g.ihy;
    const double fz = (double)p.z * g.ihz;
    const double lim = 1.0e9;
    const bool ok = fabs(fx) < lim && fabs(fy) < lim && fabs(fz) < lim && isfinite((double)p.w);
    if (ok) {
      const int ox = wrapi((int)round(fx) - g.nx, g.X);
      const int oy = wrapi((int)round(fy) - g.ny, g.Y);
      const int oz = wrapi((int)round(fz) - g.nz, g.Z);
      const BinShape& b = S.b;
      const int qx = ox / b.Bx, qy = oy / b.By, qz = oz / b.Bz;
      const int bin = qx + b.nbx * (qy + b.nby * qz);
      const int slot = atomicAdd(&d.counts[S.bofs + bin], 1);
      if (slot < S.cap_bin) {
        const long long pos = S.slot0 + (long long)bin * S.cap_bin + slot;
        ((V4<CT>*)d.sorted)[pos] = p;
        d.perm[pos] = i;
        d.loc[pos] = (ox - qx * b.Bx) | ((oy - qy * b.By) << 8) | ((oz - qz * b.Bz) << 16);
      } else {
        d.ovf[atomicAdd(d.novf, 1)] = i;
      }
      aq = fabs((double)p.w);
    } else {
      atomicAdd(&d.stats[s].invalid, 1);
    }
  }
  // max |q| per system: one atomic per wave when the wave holds one system
  const int s0 = __builtin_amdgcn_readfirstlane(s);
  const bool uniform = __all(s == s0);
  if (uniform) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) aq = fmax(aq, __shfl_xor(aq, m, kWave));
    if (lane == 0 && s0 >= 0 && aq > 0.0)
      atomicMax(&d.stats[s0].qmax_bits, (unsigned long long)__double_as_longlong(aq));
  } else if (s >= 0 && aq > 0.0) {
    atomicMax(&d.stats[s].qmax_bits, (unsigned long long)__double_as_longlong(aq));
  }
}

// The batched binning with slot claims aggregated per workgroup (as
// k_bin_agg): 1024 consecutive particles span few systems, so the bins they
// touch lie in a short range of the concatenated bin numbering; each touched
// bin is claimed with one global atomic.  A workgroup spanning 32 systems or
// more, or a bin range over kBatchAggBins, claims one slot per particle.
constexpr int kBatchAggBins = 8192;
template <typename CT>
__global__ void __launch_bounds__(1024) k_batch_bin_agg(BatchDev d, const V4<CT>* __restrict__ xyzq) {
  __shared__ int hist[kBatchAggBins];
  const int i0 = blockIdx.x * 1024;
  const int i = i0 + (int)threadIdx.x;
  const int lane = threadIdx.x & 63;
  // systems sa..sb hold this workgroup's particles; their bins (numbered in
  // the internal system order) lie in [base, base + range)
  const int sa = batch_system(d.pofs, d.nsys, i0);
  const int sb = batch_system(d.pofs, d.nsys, min(d.ntot, i0 + 1024) - 1);
  int base = 0, range = kBatchAggBins + 1;
  if (sb - sa < 32) {
    int lo = d.sys[sa].bofs, hi = lo + d.sys[sa].b.nbins;
    for (int t = sa + 1; t <= sb; t++) {
      lo = min(lo, d.sys[t].bofs);
      hi = max(hi, d.sys[t].bofs + d.sys[t].b.nbins);
    }
    base = lo;
    range = hi - lo;
  }
  const bool agg = range <= kBatchAggBins;
  if (agg)
    for (int j = threadIdx.x; j < range; j += 1024) hist[j] = 0;
  __syncthreads();
  int s = -1, gb = -1, rank = 0, lc = 0;
  double aq = 0.0;
  V4<CT> p{0, 0, 0, 0};
  if (i < d.ntot) {
    s = batch_system(d.pofs, d.nsys, i);
    const BatchSys& S = d.sys[s];
    const Geom& g = S.g;
    p = xyzq[i];
    const double fx = (double)p.x * g.ihx;
    const double fy = (double)p.y * g.ihy;
    const double fz = (double)p.z * g.ihz;
    const double lim = 1.0e9;
    const bool ok = fabs(fx) < lim && fabs(fy) < lim && fabs(fz) < lim && isfinite((double)p.w);
    if (ok) {
      const int ox = wrapi((int)round(fx) - g.nx, g.X);
      const int oy = wrapi((int)round(fy) - g.ny, g.Y);
      const int oz = wrapi((int)round(fz) - g.nz, g.Z);
      const BinShape& b = S.b;
      const int qx = ox / b.Bx, qy = oy / b.By, qz = oz / b.Bz;
      gb = S.bofs + qx + b.nbx * (qy + b.nby * qz);
      lc = (ox - qx * b.Bx) | ((oy - qy * b.By) << 8) | ((oz - qz * b.Bz) << 16);
      if (agg) rank = atomicAdd(&hist[gb - base], 1);
      aq = fabs((double)p.w);
    } else {
      atomicAdd(&d.stats[s].invalid, 1);
    }
  }
  __syncthreads();
  if (agg)
    for (int j = threadIdx.x; j < range; j += 1024) {
      const int c = hist[j];
      if (c > 0) hist[j] = atomicAdd(&d.counts[base + j], c);
    }
  __syncthreads();
  if (gb >= 0) {
    const BatchSys& S = d.sys[s];
    const int slot = agg ? hist[gb - base] + rank : atomicAdd(&d.counts[gb], 1);
    if (slot < S.cap_bin) {
      const long long pos = S.slot0 + (long long)(gb - S.bofs) * S.cap_bin + slot;
      ((V4<CT>*)d.sorted)[pos] = p;
      d.perm[pos] = i;
      d.loc[pos] = lc;
    } else {
      d.ovf[atomicAdd(d.novf, 1)] = i;
    }
  }
  // max |q| per system: one atomic per wave when the wave holds one system
  const int s0 = __builtin_amdgcn_readfirstlane(s);
  const bool uniform = __all(s == s0);
  if (uniform) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) aq = fmax(aq, __shfl_xor(aq, m, kWave));
    if (lane == 0 && s0 >= 0 && aq > 0.0)
      atomicMax(&d.stats[s0].qmax_bits, (unsigned long long)__double_as_longlong(aq));
  } else if (s >= 0 && aq > 0.0) {
    atomicMax(&d.stats[s].qmax_bits, (unsigned long long)__double_as_longlong(aq));
  }
}

// sysp and items are read-only for the whole launch (separate restrict
// arguments, so their uniform loads go to scalar registers)
template <typename CT, int NZC, int BS, typename ACC>
__global__ void __launch_bounds__(BS)
    k_batch_spread(BatchDev d, const BatchSys* __restrict__ sysp, const int2* __restrict__ items,
                   int nitems) {
  extern __shared__ __align__(32) unsigned char spread_smem[];
  const ItemWalk w = item_walk(nitems);
  for (int it = w.it; it < w.end; it += w.step) {
    const int2 item = items[it];
    const BatchSys& S = sysp[item.x];
    const int bin = item.y;
    const int call = d.counts[S.bofs + bin];
    const int cnt = min(call, S.cap_bin);
    if (cnt == 0) {
      if (d.bricks && threadIdx.x == 0) d.bshift[S.bofs + bin] = -1;  // no brick this step
      continue;
    }
    const BinStats st = d.stats[item.x];
    const int eg = fixed_exponent(st.qmax_bits, S.n - st.invalid, S.images, S.pref_bound);
    const SpreadArgs a = batch_spread_args(S);
    const long long s0 = S.slot0 + (long long)bin * S.cap_bin;
    const long long vol = (long long)S.b.Lx * S.b.Ly * S.b.Lz;
    spread_bin<CT, NZC, BS, ACC>(a, bin, cnt, call, eg, st.qmax_bits, (const V4<CT>*)d.sorted + s0,
                                 d.loc + s0, d.grid64 + S.gofs, spread_smem,
                                 d.bricks ? (ACC*)d.bricks + S.brick0 + bin * vol : nullptr,
                                 d.bshift + S.bofs);
  }
}

// overflow particles of all systems, one wave each
template <typename CT, typename ACC>
__global__ void __launch_bounds__(256)
    k_batch_spread_ovf(BatchDev d, const V4<CT>* __restrict__ xyzq) {
  const int novf = *d.novf;
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int k = gw; k < novf; k += nw) {
    const int i = d.ovf[k];
    const int s = batch_system(d.pofs, d.nsys, i);
    const BatchSys& S = d.sys[s];
    const BinStats st = d.stats[s];
    const int eg = fixed_exponent(st.qmax_bits, S.n - st.invalid, S.images, S.pref_bound);
    spread_overflow<CT, ACC>(batch_spread_args(S), eg, st.qmax_bits, d.counts + S.bofs, xyzq[i],
                             d.grid64 + S.gofs);
  }
}

// brick-store flush of the batch: one workgroup per (system, bin) item
template <typename T, typename ACC, bool TILE4>
__global__ void __launch_bounds__(TILE4 ? 128 : 512)
    k_batch_reduce(BatchDev d, const BatchSys* __restrict__ sysp, const int2* __restrict__ items,
                   T* __restrict__ rho) {
  const int2 item = items[blockIdx.x];
  const BatchSys& S = sysp[item.x];
  const BinStats st = d.stats[item.x];
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, S.n - st.invalid, S.images, S.pref_bound));
  const bool ovf = *d.novf > 0;
  const ACC* bricks = (const ACC*)d.bricks + S.brick0;
  if constexpr (TILE4)
    reduce_tile4_bin<T, false>(S.g, S.b, item.y, (const unsigned*)bricks, d.bshift + S.bofs, inv, ovf,
                        d.grid64 + S.gofs, rho + S.gofs);
  else
    reduce_tile_bin<T, ACC>(S.g, S.b, item.y, bricks, d.bshift + S.bofs, inv, ovf,
                            d.grid64 + S.gofs, rho + S.gofs);
}

// int64 -> T for every system's grid (one chunk per workgroup), int64 cleared
template <typename T>
__global__ void __launch_bounds__(256)
    k_batch_convert(BatchDev d, const BatchChunk* __restrict__ chunks, T* __restrict__ rho) {
  const BatchChunk c = chunks[blockIdx.x];
  const BatchSys& S = d.sys[c.sys];
  const BinStats st = d.stats[c.sys];
  const double inv =
      ldexp(1.0, -fixed_exponent(st.qmax_bits, S.n - st.invalid, S.images, S.pref_bound));
  unsigned long long* g = d.grid64 + S.gofs + c.start;
  T* out = rho + S.gofs + c.start;
  for (int i = threadIdx.x; i < c.count; i += blockDim.x) {
    out[i] = (T)((double)(long long)g[i] * inv);
    g[i] = 0ull;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
Geom make_geom(int X, int Y, int Z, double bx, double by, double bz, double rcut) {
  Geom g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.hx = bx / (double)X;
  g.hy = by / (double)Y;
  g.hz = bz / (double)Z;
  g.ihx = 1.0 / g.hx;
  g.ihy = 1.0 / g.hy;
  g.ihz = 1.0 / g.hz;
  g.nx = (int)ceil(rcut / g.hx);
  g.ny = (int)ceil(rcut / g.hy);
  g.nz = (int)ceil(rcut / g.hz);
  return g;
}

// The LDS path needs the 3(2n+1) factors in one wave, at most kMaxK column
// passes, brick rows no longer than a wave and a brick within the budget.
BinShape choose_bins(const Geom& g, size_t elem, size_t budget, int pref_edge) {
  BinShape b;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const bool lanes_ok = NX + NY + NZ <= 64 && (NX * NY + 63) / 64 <= kMaxK;
  int B[3] = {pref_edge, pref_edge, pref_edge};
  const int M[3] = {g.X, g.Y, g.Z};
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int a = 0; a < 3; a++) {
    if (B[a] > M[a]) B[a] = M[a];
    if (B[a] + 2 * n[a] > 64) B[a] = 64 - 2 * n[a];
    if (B[a] < 1) B[a] = 0;
  }
  bool fits = lanes_ok && B[0] > 0 && B[1] > 0 && B[2] > 0;
  // shrink the largest edge until the brick fits
  while (fits) {
    size_t bytes = (size_t)(B[0] + 2 * n[0]) * (B[1] + 2 * n[1]) * (B[2] + 2 * n[2]) * elem;
    if (bytes <= budget) break;
    int a = 0;
    for (int k = 1; k < 3; k++)
      if (B[k] > B[a]) a = k;
    if (B[a] == 1) {
      fits = false;
      break;
    }
    B[a] = (B[a] + 1) / 2;
  }
  b.direct = 0;
  if (!fits) {  // one bin holding every particle, global-memory kernels
    b.Bx = g.X;
    b.By = g.Y;
    b.Bz = g.Z;
    b.nbx = b.nby = b.nbz = 1;
    b.nbins = 1;
    b.Lx = b.Ly = b.Lz = 0;
    b.direct = 1;
    return b;
  }
  b.Bx = B[0];
  b.By = B[1];
  b.Bz = B[2];
  b.nbx = (g.X + B[0] - 1) / B[0];
  b.nby = (g.Y + B[1] - 1) / B[1];
  b.nbz = (g.Z + B[2] - 1) / B[2];
  b.nbins = b.nbx * b.nby * b.nbz;
  b.Lx = B[0] + 2 * g.nx;
  b.Ly = B[1] + 2 * g.ny;
  b.Lz = B[2] + 2 * g.nz;
  return b;
}

hipError_t launch_bin_clear(int* counts, int nbins, BinStats* stats, hipStream_t s) {
  k_bin_clear<<<grid_for(nbins, 256), 256, 0, s>>>(counts, nbins, stats);
  return hipGetLastError();
}

hipError_t launch_bin(bool dbl, const Geom& g, const BinShape& b, Binning& w, hipStream_t s) {
  // bin_agg: particles per workgroup / 256
  const int agg = tuning().bin_agg;
  if (w.n > 0 && b.nbins <= kBinAggMaxBins && agg > 0) {
    const int per_wg = 256 * agg;
    const unsigned grid = (unsigned)((w.n + per_wg - 1) / per_wg);
#define GSE_BIN_AGG(TT, BT, P)                                                                   \
  k_bin_agg<TT, BT, P><<<grid, BT, 0, s>>>(g, b, w.n, w.cap_bin, (const V4<TT>*)w.xyzq, w.counts, \
                                           w.stats, w.ovf, w.perm, w.loc, (V4<TT>*)w.sorted)
#define GSE_BIN_AGG_T(TT)                  \
  if (agg == 1)                            \
    GSE_BIN_AGG(TT, 256, 1);               \
  else if (agg == 2)                       \
    GSE_BIN_AGG(TT, 512, 1);               \
  else if (agg == 4)                       \
    GSE_BIN_AGG(TT, 1024, 1);              \
  else                                     \
    GSE_BIN_AGG(TT, 1024, 2);
    if (dbl) {
      GSE_BIN_AGG_T(double)
    } else {
      GSE_BIN_AGG_T(float)
    }
#undef GSE_BIN_AGG_T
#undef GSE_BIN_AGG
    return hipGetLastError();
  }
  if (w.n > 0 && w.scan && b.nbins > kBinAggMaxBins && agg > 0 &&
      super_shape(w.n, b).nsuper <= kSuperMax) {
    const SuperShape q = super_shape(w.n, b);
    const int cb = tuning().bin_coarse_block;
    const size_t slots = (size_t)q.nsuper * kStripes * q.cap;
    unsigned* scount = (unsigned*)w.scan;
    char* sxp = (char*)w.scan + super_count_bytes(q);
    int* sidx = (int*)(sxp + slots * (dbl ? 32 : 16));
    int* sloc = sidx + slots;
    int* spill = sloc + slots;
#define GSE_BIN_COARSE(TT, BT)                                                              \
  k_bin_coarse<TT, BT><<<(w.n + BT * kBinSortPPT - 1) / (BT * kBinSortPPT), BT,              \
                         coarse_lds_bytes(q.nsuper, BT * kBinSortPPT, sizeof(TT)), s>>>(     \
      g, b, q, w.n, (const V4<TT>*)w.xyzq, scount, (V4<TT>*)sxp, sidx, sloc, spill, w.stats)
    if (dbl) {
      GSE_BIN_COARSE(double, 512);  // 32-byte 4-vectors: half the staging
      k_bin_fine<double><<<q.nsuper, kBinSortThreads, 0, s>>>(
          g, b, q, w.cap_bin, (const V4<double>*)w.xyzq, scount, (const V4<double>*)sxp, sidx,
          sloc, spill, w.stats, w.ovf, w.counts, w.perm, w.loc, (V4<double>*)w.sorted);
    } else {
      if (cb == 256)
        GSE_BIN_COARSE(float, 256);
      else if (cb == 512)
        GSE_BIN_COARSE(float, 512);
      else
        GSE_BIN_COARSE(float, 1024);
      k_bin_fine<float><<<q.nsuper, kBinSortThreads, 0, s>>>(
          g, b, q, w.cap_bin, (const V4<float>*)w.xyzq, scount, (const V4<float>*)sxp, sidx,
          sloc, spill, w.stats, w.ovf, w.counts, w.perm, w.loc, (V4<float>*)w.sorted);
    }
#undef GSE_BIN_COARSE
    return hipGetLastError();
  }
  if (w.n > 0) {
    const unsigned grid = (unsigned)((w.n + 255) / 256);
    if (dbl)
      k_bin<double><<<grid, 256, 0, s>>>(g, b, w.n, w.cap_bin, (const V4<double>*)w.xyzq,
                                         w.counts, w.stats, w.ovf, w.perm, w.loc,
                                         (V4<double>*)w.sorted);
    else
      k_bin<float><<<grid, 256, 0, s>>>(g, b, w.n, w.cap_bin, (const V4<float>*)w.xyzq, w.counts,
                                        w.stats, w.ovf, w.perm, w.loc, (V4<float>*)w.sorted);
  }
  return hipGetLastError();
}

template <typename CT>
static hipError_t spread_t(const Geom& g, const BinShape& b, double sigma, bool normalize,
                           Binning& w, unsigned long long* grid64, hipStream_t s) {
  SpreadArgs a;
  a.g = g;
  a.b = b;
  a.inv2s2 = 1.0 / (2.0 * sigma * sigma);
  a.pref = pow(2.0 * M_PI * sigma * sigma, -3.0 / 2.0);
  a.pref1 = sqrt(2.0 * M_PI * sigma * sigma);
  a.normalize = normalize ? 1 : 0;
  a.cap_bin = w.cap_bin;
  a.n = w.n;
  a.images = w.images;
  a.pref_bound = w.pref_bound;
  a.order = w.order;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  if (b.direct) {
    a.cap = 0;
    a.nwg_bins = 1;
    k_spread_direct<CT><<<2048, 256, 0, s>>>(a, (const V4<CT>*)w.sorted, w.counts, w.counts_other,
                                             w.stats, w.stats_other, grid64);
    return hipGetLastError();
  }
  const Tuning& tu = tuning();
  if (tu.spread_cols && b.Bz <= 8 && b.Lx * b.Ly <= 1024 && (NZ == 11 || NZ == 13 || NZ == 15 ||
                                                            NZ == 17)) {
    const bool acc32 = sizeof(CT) == 4 && tu.spread_lds32;
    int cap = tu.spread_stage / (int)(4 * sizeof(CT) + 8 * (NX + NY + NZ) + 8 + 4);
    if (cap > 256) cap = 256;
    if (cap < 8) cap = 8;
    a.cap = cap;
    const size_t lds = spread_lds(0, cap, NX + NY + NZ, sizeof(CT)).total;
    a.nwg_bins = ((b.nbins + 7) / 8) * 8;
    const unsigned nwg = (unsigned)(a.nwg_bins + kOverflowWG);
    const int cols = b.Lx * b.Ly;
    const int bs = cols <= 256 ? 256 : cols <= 512 ? 512 : 1024;
    const V4<CT>* xq = (const V4<CT>*)w.xyzq;
    const V4<CT>* srt = (const V4<CT>*)w.sorted;
#define GSE_COLS_LAUNCH(NZV, BSV, ACCT)                                                       \
  k_spread_cols<CT, NZV, 8, BSV, ACCT><<<nwg, BSV, lds, s>>>(a, xq, srt, w.counts,            \
                                                             w.counts_other, w.stats,         \
                                                             w.stats_other, w.ovf, grid64)
#define GSE_COLS_BS(NZV, ACCT)            \
  if (bs == 256)                          \
    GSE_COLS_LAUNCH(NZV, 256, ACCT);      \
  else if (bs == 512)                     \
    GSE_COLS_LAUNCH(NZV, 512, ACCT);      \
  else                                    \
    GSE_COLS_LAUNCH(NZV, 1024, ACCT);
#define GSE_COLS_NZ(ACCT)                      \
  switch (NZ) {                                \
    case 11: GSE_COLS_BS(11, ACCT) break;      \
    case 13: GSE_COLS_BS(13, ACCT) break;      \
    case 15: GSE_COLS_BS(15, ACCT) break;      \
    default: GSE_COLS_BS(17, ACCT) break;      \
  }
    if (acc32) {
      GSE_COLS_NZ(unsigned)
    } else {
      GSE_COLS_NZ(unsigned long long)
    }
#undef GSE_COLS_NZ
#undef GSE_COLS_BS
#undef GSE_COLS_LAUNCH
    return hipGetLastError();
  }
  // stage as many particles as fit beside the brick in the workgroup's LDS
  const bool acc32 = sizeof(CT) == 4 && tu.spread_lds32;
  if (tu.spread_mma && spread_mma_ok(sizeof(CT) == 8, g, b, normalize, acc32, w.bricks != nullptr)) {
    a.cap = kMmaK;
    a.nwg_bins = ((b.nbins + 7) / 8) * 8;
    k_spread_mma<<<(unsigned)a.nwg_bins, kMmaThreads, kMmaLds, s>>>(
        a, (const V4<float>*)w.xyzq, (const V4<float>*)w.sorted, w.counts, w.counts_other,
        w.stats, w.stats_other, w.ovf, grid64, (unsigned*)w.bricks, w.bshift, w.loc);
    return hipGetLastError();
  }
  const int vol = b.Lx * b.Ly * b.Lz;
  const int vol_eff = vol * (acc32 ? 4 : 8) / 8 + 1;  // brick size in 8-byte units
  const size_t lds_total = (size_t)vol_eff * 8 + (size_t)tu.spread_stage;
  int cap = 256;
  while (cap > 1 && spread_lds(vol_eff, cap, NX + NY + NZ, sizeof(CT)).total > lds_total) cap--;
  a.cap = cap;
  const size_t lds = spread_lds(vol_eff, cap, NX + NY + NZ, sizeof(CT)).total;
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  a.nwg_bins = ((b.nbins + 7) / 8) * 8;
  const unsigned nwg = (unsigned)(a.nwg_bins + kOverflowWG);
  const V4<CT>* xq = (const V4<CT>*)w.xyzq;
  const V4<CT>* srt = (const V4<CT>*)w.sorted;
  const int bs = tu.spread_block ? tu.spread_block : (sizeof(CT) == 4 ? 512 : 256);
#define GSE_SPREAD_LAUNCH(NZV, BSV, ACCT)                                                    \
  k_spread<CT, NZV, BSV, ACCT><<<nwg, BSV, lds, s>>>(a, xq, srt, w.counts, w.counts_other,  \
                                                     w.stats, w.stats_other, w.ovf, grid64, \
                                                     (ACCT*)w.bricks, w.bshift, w.loc)
#define GSE_SPREAD_BS(NZV, ACCT)                      \
  if (bs == 512)                                      \
    GSE_SPREAD_LAUNCH(NZV, 512, ACCT);                \
  else if (bs == 1024)                                \
    GSE_SPREAD_LAUNCH(NZV, 1024, ACCT);               \
  else                                                \
    GSE_SPREAD_LAUNCH(NZV, 256, ACCT);
#define GSE_SPREAD_NZ(ACCT)                           \
  switch (NZ) {                                       \
    case 11: GSE_SPREAD_BS(11, ACCT) break;           \
    case 13:                                          \
      if (bs == 512 && b.Lx * b.Ly == 400) {          \
        k_spread<CT, 13, 512, ACCT, 400><<<nwg, 512, lds, s>>>(                             \
            a, xq, srt, w.counts, w.counts_other, w.stats, w.stats_other, w.ovf, grid64,     \
            (ACCT*)w.bricks, w.bshift, w.loc);        \
        break;                                        \
      }                                               \
      if constexpr (sizeof(CT) == 8) { /* FP64: 256 threads */                               \
        if (bs == 256 && b.Lx * b.Ly == 400) {                                               \
          k_spread<CT, 13, 256, ACCT, 400><<<nwg, 256, lds, s>>>(                           \
              a, xq, srt, w.counts, w.counts_other, w.stats, w.stats_other, w.ovf, grid64,   \
              (ACCT*)w.bricks, w.bshift, w.loc);      \
          break;                                      \
        }                                             \
      }                                               \
      GSE_SPREAD_BS(13, ACCT) break;                  \
    case 15: GSE_SPREAD_BS(15, ACCT) break;           \
    case 17: GSE_SPREAD_BS(17, ACCT) break;           \
    default: GSE_SPREAD_BS(0, ACCT) break;            \
  }
  if (acc32) {
    GSE_SPREAD_NZ(unsigned)
  } else {
    GSE_SPREAD_NZ(unsigned long long)
  }
#undef GSE_SPREAD_NZ
#undef GSE_SPREAD_BS
#undef GSE_SPREAD_LAUNCH
  return hipGetLastError();
}

hipError_t launch_spread(bool dbl, const Geom& g, const BinShape& b, double sigma,
                         bool normalize, Binning& w, unsigned long long* grid64,
                         hipStream_t s) {
  return dbl ? spread_t<double>(g, b, sigma, normalize, w, grid64, s)
             : spread_t<float>(g, b, sigma, normalize, w, grid64, s);
}

double spread_weight_bound(const Geom& g, double sigma, bool normalize) {
  if (!normalize) return 1.0;
  const double h[3] = {g.hx, g.hy, g.hz};
  double f = 1.0;
  for (int a = 0; a < 3; a++) {
    double w = sqrt(2.0 * M_PI) * sigma * exp(h[a] * h[a] / (8.0 * sigma * sigma));
    f *= w > 1.0 ? w : 1.0;
  }
  return f;
}

hipError_t launch_fixed_to_grid(bool dbl, const unsigned long long* grid64, void* out,
                                size_t count, const Binning& w, bool clear, hipStream_t s) {
  const unsigned grid = grid_for(count, 256, 8192);
  if (dbl)
    k_fixed_to_grid<double><<<grid, 256, 0, s>>>((unsigned long long*)grid64, (double*)out,
                                                 count, w.stats, w.n, w.images, w.pref_bound,
                                                 clear ? 1 : 0);
  else
    k_fixed_to_grid<float><<<grid, 256, 0, s>>>((unsigned long long*)grid64, (float*)out, count,
                                                w.stats, w.n, w.images, w.pref_bound,
                                                clear ? 1 : 0);
  return hipGetLastError();
}

static int brick_axis_max(int M, int B, int nb, int L) {
  int worst = 0;
  for (int p = 0; p < M; p++) {
    int c = 0;
    for (int q = p; q - (nb - 1) * B < L; q += M) {
      int blo = q - L + 1;
      blo = blo <= 0 ? 0 : (blo + B - 1) / B;
      const int bhi = std::min(q / B, nb - 1);
      if (bhi >= blo) c += bhi - blo + 1;
    }
    worst = std::max(worst, c);
  }
  return worst;
}

bool brick_reduce_ok(const Geom& g, const BinShape& b) {
  if (b.direct) return false;
  return brick_axis_max(g.X, b.Bx, b.nbx, b.Lx) <= KR &&
         brick_axis_max(g.Y, b.By, b.nby, b.Ly) <= KR &&
         brick_axis_max(g.Z, b.Bz, b.nbz, b.Lz) <= KR;
}

bool brick_tile_ok(const Geom& g, const BinShape& b) {
  return b.Bx == 8 && b.By == 8 && b.Bz == 8 && g.X % 8 == 0 && g.Y % 8 == 0 && g.Z % 8 == 0 &&
         b.nbx >= 3 && b.nby >= 3 && b.nbz >= 3 && b.Lx <= 24 && b.Ly <= 24 && b.Lz <= 24;
}

int brick_reduce_path(bool dbl, bool acc32, const Geom& g, const BinShape& b) {
  if (!(brick_tile_ok(g, b) && tuning().reduce_tile)) return kReduceGeneric;
  if (!dbl && acc32 && b.Lx % 4 == 0) return tuning().reduce_nt ? kReduceTile4NT : kReduceTile4;
  if (!acc32 && b.Lx % 2 == 0 && tuning().reduce_tile2) return kReduceTile2;
  return kReduceTile;
}

hipError_t launch_brick_reduce(bool dbl, bool acc32, const Geom& g, const BinShape& b,
                               const Binning& w, unsigned long long* grid64, void* out,
                               hipStream_t s) {
  const int path = brick_reduce_path(dbl, acc32, g, b);
  if (path != kReduceGeneric) {
    const int nwg = (b.nbins + 7) / 8 * 8;
    if (path == kReduceTile4 || path == kReduceTile4NT) {
      auto k = path == kReduceTile4NT ? k_brick_reduce_tile4<float, true>
                                      : k_brick_reduce_tile4<float, false>;
      k<<<nwg, 128, 0, s>>>(g, b, nwg, (const unsigned*)w.bricks, w.bshift, w.stats, w.n,
                            w.images, w.pref_bound, grid64, (float*)out);
      return hipGetLastError();
    }
    if (path == kReduceTile2) {
      if (dbl)
        k_brick_reduce_tile2<double><<<nwg, 256, 0, s>>>(
            g, b, nwg, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
            w.pref_bound, grid64, (double*)out);
      else
        k_brick_reduce_tile2<float><<<nwg, 256, 0, s>>>(
            g, b, nwg, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
            w.pref_bound, grid64, (float*)out);
      return hipGetLastError();
    }
    if (dbl)
      k_brick_reduce_tile<double, unsigned long long><<<nwg, 512, 0, s>>>(
          g, b, nwg, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
          w.pref_bound, grid64, (double*)out);
    else if (acc32)
      k_brick_reduce_tile<float, unsigned><<<nwg, 512, 0, s>>>(
          g, b, nwg, (const unsigned*)w.bricks, w.bshift, w.stats, w.n, w.images, w.pref_bound,
          grid64, (float*)out);
    else
      k_brick_reduce_tile<float, unsigned long long><<<nwg, 512, 0, s>>>(
          g, b, nwg, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
          w.pref_bound, grid64, (float*)out);
    return hipGetLastError();
  }
  const size_t np = (size_t)g.X * g.Y * g.Z;
  const unsigned grid = grid_for(np, 256, 8192);
  if (dbl)
    k_brick_reduce<double, unsigned long long><<<grid, 256, 0, s>>>(
        g, b, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
        w.pref_bound, grid64, (double*)out);
  else if (acc32)
    k_brick_reduce<float, unsigned><<<grid, 256, 0, s>>>(g, b, (const unsigned*)w.bricks,
                                                          w.bshift, w.stats, w.n, w.images,
                                                          w.pref_bound, grid64, (float*)out);
  else
    k_brick_reduce<float, unsigned long long><<<grid, 256, 0, s>>>(
        g, b, (const unsigned long long*)w.bricks, w.bshift, w.stats, w.n, w.images,
        w.pref_bound, grid64, (float*)out);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, size_t count, hipStream_t s) {
  k_fill_u64<<<grid_for(count, 256, 8192), 256, 0, s>>>(p, count);
  return hipGetLastError();
}

hipError_t launch_fill(bool dbl, void* p, size_t count, hipStream_t s) {
  if (dbl)
    k_fill<double><<<grid_for(count, 256, 8192), 256, 0, s>>>((double*)p, count);
  else
    k_fill<float><<<grid_for(count, 256, 8192), 256, 0, s>>>((float*)p, count);
  return hipGetLastError();
}


size_t spread_lds_bytes(bool dbl, const Geom& g, const BinShape& b, int* cap_out) {
  const Tuning& tu = tuning();
  const bool acc32 = !dbl && tu.spread_lds32;
  const int NW = 2 * (g.nx + g.ny + g.nz) + 3;
  const int vol = b.Lx * b.Ly * b.Lz;
  const int vol_eff = vol * (acc32 ? 4 : 8) / 8 + 1;
  const size_t lds_total = (size_t)vol_eff * 8 + (size_t)tu.spread_stage;
  const int ct = dbl ? 8 : 4;
  int cap = 256;
  while (cap > 1 && spread_lds(vol_eff, cap, NW, ct).total > lds_total) cap--;
  if (cap_out) *cap_out = cap;
  return spread_lds(vol_eff, cap, NW, ct).total;
}

hipError_t launch_batch_clear(const BatchDev& d, hipStream_t s) {
  const int n = d.total_bins > d.nsys ? d.total_bins : d.nsys;
  k_batch_clear<<<grid_for(n, 256), 256, 0, s>>>(d.counts, d.total_bins, d.stats, d.nsys, d.novf);
  return hipGetLastError();
}

hipError_t launch_batch_bin(bool dbl, const BatchDev& d, const void* xyzq, hipStream_t s) {
  if (d.ntot <= 0) return hipSuccess;
  if (tuning().bin_agg) {
    const unsigned grid = (unsigned)((d.ntot + 1023) / 1024);
    if (dbl)
      k_batch_bin_agg<double><<<grid, 1024, 0, s>>>(d, (const V4<double>*)xyzq);
    else
      k_batch_bin_agg<float><<<grid, 1024, 0, s>>>(d, (const V4<float>*)xyzq);
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)((d.ntot + 255) / 256);
  if (dbl)
    k_batch_bin<double><<<grid, 256, 0, s>>>(d, (const V4<double>*)xyzq);
  else
    k_batch_bin<float><<<grid, 256, 0, s>>>(d, (const V4<float>*)xyzq);
  return hipGetLastError();
}

template <typename CT, typename ACC>
static hipError_t batch_spread_t(int nz, int bs, size_t lds, int grid, const BatchDev& d,
                                 const int2* items, int nitems, hipStream_t s) {
  if (nitems > 0) {
    const unsigned g = (unsigned)grid;
#define GSE_BS_LAUNCH(NZV, BSV) \
  k_batch_spread<CT, NZV, BSV, ACC><<<g, BSV, lds, s>>>(d, d.sys, items, nitems)
#define GSE_BS_BS(NZV)          \
  if (bs == 512)                \
    GSE_BS_LAUNCH(NZV, 512);    \
  else                          \
    GSE_BS_LAUNCH(NZV, 256);
    switch (nz) {
      case 13: GSE_BS_BS(13) break;
      case 15: GSE_BS_BS(15) break;
      case 17: GSE_BS_BS(17) break;
      default: GSE_BS_BS(0) break;
    }
#undef GSE_BS_BS
#undef GSE_BS_LAUNCH
  }
  return hipGetLastError();
}

hipError_t launch_batch_spread(bool dbl, int nz, int block, size_t lds, int grid,
                               const BatchDev& d, const int2* items, int nitems, hipStream_t s) {
  if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
  if (dbl) return batch_spread_t<double, unsigned long long>(nz, block, lds, grid, d, items,
                                                             nitems, s);
  if (tuning().spread_lds32)
    return batch_spread_t<float, unsigned>(nz, block, lds, grid, d, items, nitems, s);
  return batch_spread_t<float, unsigned long long>(nz, block, lds, grid, d, items, nitems, s);
}

hipError_t launch_batch_spread_overflow(bool dbl, const BatchDev& d, const void* xyzq,
                                        hipStream_t s) {
  if (dbl)
    k_batch_spread_ovf<double, unsigned long long><<<64, 256, 0, s>>>(d, (const V4<double>*)xyzq);
  else if (tuning().spread_lds32)
    k_batch_spread_ovf<float, unsigned><<<64, 256, 0, s>>>(d, (const V4<float>*)xyzq);
  else
    k_batch_spread_ovf<float, unsigned long long><<<64, 256, 0, s>>>(d, (const V4<float>*)xyzq);
  return hipGetLastError();
}

hipError_t launch_batch_reduce(bool dbl, bool acc32, bool tile4, const BatchDev& d,
                               const int2* items, int nitems, void* rho, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (dbl)
    k_batch_reduce<double, unsigned long long, false><<<nitems, 512, 0, s>>>(d, d.sys, items,
                                                                             (double*)rho);
  else if (acc32 && tile4)
    k_batch_reduce<float, unsigned, true><<<nitems, 128, 0, s>>>(d, d.sys, items, (float*)rho);
  else if (acc32)
    k_batch_reduce<float, unsigned, false><<<nitems, 512, 0, s>>>(d, d.sys, items, (float*)rho);
  else
    k_batch_reduce<float, unsigned long long, false><<<nitems, 512, 0, s>>>(d, d.sys, items,
                                                                            (float*)rho);
  return hipGetLastError();
}

hipError_t launch_batch_convert(bool dbl, const BatchDev& d, const BatchChunk* chunks,
                                int nchunks, void* rho, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  if (dbl)
    k_batch_convert<double><<<nchunks, 256, 0, s>>>(d, chunks, (double*)rho);
  else
    k_batch_convert<float><<<nchunks, 256, 0, s>>>(d, chunks, (float*)rho);
  return hipGetLastError();
}

}  // namespace gse
