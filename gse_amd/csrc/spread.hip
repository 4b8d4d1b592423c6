// Cell-binned Gaussian charge spreading with fixed-point int64 accumulation
// (CpuGaussCharge::spreadChargeToGrid, CpuGaussCharge.cpp:235-384).
//
// Pipeline (all on one stream):
//   k_bin_reset    zero bin counters and the fixed-point statistics
//   k_bin_count    cube origin o = wrap(round(x*inv_h) - n) per particle,
//                  bin of o, slot in the bin, max|q|
//   k_bin_scan     exclusive scan of the counters (one workgroup) and the
//                  fixed-point scale 2^e chosen from N, max|q| and pref
//   k_bin_scatter  particles copied into bin order (sorted 4-vectors + perm)
//   k_spread       one workgroup per bin: the (B+2n)^3 brick lives in LDS as
//                  int64; every wave takes one particle at a time, lanes own
//                  (tx,ty) columns of its cube and walk tz, adding
//                  round(q*pref*wx*wy*wz*2^e) with ds_add_u64; the brick is
//                  then added into the global int64 grid with 64-bit atomics.
//   k_fixed_to_grid  int64 -> T with 2^-e, and the int64 grid re-zeroed.
// Integer addition is associative, so the grid is bit-identical run to run
// whatever order the particles and workgroups run in.
#include <math.h>

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

__global__ void k_bin_reset(int* counts, int n1, FixedParams* fp) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n1; i += gridDim.x * blockDim.x)
    counts[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    fp->qmax_bits = 0ull;
    fp->invalid = 0;
  }
}

template <typename CT>
__global__ void k_bin_count(Geom g, BinShape b, int n, const V4<CT>* __restrict__ xyzq,
                            int* __restrict__ counts, int* __restrict__ atom_bin,
                            int* __restrict__ atom_slot, FixedParams* fp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double aq = 0.0;
  if (i < n) {
    V4<CT> p = xyzq[i];
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
      const int bin = ox / b.Bx + b.nbx * (oy / b.By + b.nby * (oz / b.Bz));
      atom_bin[i] = bin;
      atom_slot[i] = atomicAdd(&counts[bin], 1);
      aq = fabs((double)p.w);
    } else {
      atom_bin[i] = -1;
      atomicAdd(&fp->invalid, 1);
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) aq = fmax(aq, __shfl_xor(aq, m, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && aq > 0.0)
    atomicMax(&fp->qmax_bits, (unsigned long long)__double_as_longlong(aq));
}

// One workgroup of 1024 threads: exclusive scan in place, counts[nb] = total.
__global__ void __launch_bounds__(1024) k_bin_scan(int* counts, int nb, int n, int images,
                                                   double pref, FixedParams* fp) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + tid;
    int v = i < nb ? counts[i] : 0;
    // inclusive wave scan
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
      int s = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        int y = __shfl_up(s, o, kWave);
        if (lane >= o) s += y;
      }
      if (lane < 16) wsum[lane] = s;
    }
    __syncthreads();
    const int excl = carry + (wid ? wsum[wid - 1] : 0) + x - v;
    if (i < nb) counts[i] = excl;
    __syncthreads();
    if (tid == 1023) carry = excl + v;
    __syncthreads();
  }
  if (tid == 0) {
    counts[nb] = carry;
    // Fixed-point scale: every contribution |q pref w w w| <= qmax*pref must
    // stay below 2^51 (exact rounding trick) and the sum over all particles
    // and cube images reaching one point below 2^62.
    const double qmax = __longlong_as_double((long long)fp->qmax_bits);
    const double one = qmax * pref;
    const double all = one * (double)(carry > 0 ? carry : 1) * (double)images;
    int e = 60;
    if (one > 0.0) {
      int k1, k2;
      frexp(one, &k1);  // one < 2^k1
      frexp(all, &k2);
      e = min(51 - k1, 62 - k2);
    }
    e = max(-60, min(e, 200));
    fp->scale = ldexp(1.0, e);
    fp->inv_scale = ldexp(1.0, -e);
  }
}

template <typename CT>
__global__ void k_bin_scatter(int n, const V4<CT>* __restrict__ xyzq, const int* __restrict__ starts,
                              const int* __restrict__ atom_bin, const int* __restrict__ atom_slot,
                              int* __restrict__ perm, V4<CT>* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int bin = atom_bin[i];
  if (bin < 0) return;
  const int pos = starts[bin] + atom_slot[i];
  perm[pos] = i;
  sorted[pos] = xyzq[i];
}

// Per-lane role in the weight phase: lanes [0,NX) x-weights, [NX,NX+NY)
// y-weights, [NX+NY, NX+NY+NZ) z-weights.
struct LaneRole {
  int axis;  // 0,1,2 or -1
  int j;
};

__device__ __forceinline__ LaneRole lane_role(int lane, int NX, int NY, int NZ) {
  LaneRole r;
  if (lane < NX) {
    r.axis = 0;
    r.j = lane;
  } else if (lane < NX + NY) {
    r.axis = 1;
    r.j = lane - NX;
  } else if (lane < NX + NY + NZ) {
    r.axis = 2;
    r.j = lane - NX - NY;
  } else {
    r.axis = -1;
    r.j = 0;
  }
  return r;
}

// Weight of this lane's (axis, j) entry: exp(-t*t*inv_2sigmasq) with
// t = (c*h - x) + j*h (CpuGaussCharge.cpp:314-329); offsets in FP64.
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

template <typename CT, int NZC>
__global__ void __launch_bounds__(256)
    k_spread(Geom g, BinShape b, double inv2s2, double pref, int normalize, double pref1,
             const V4<CT>* __restrict__ sorted, const int* __restrict__ starts,
             const FixedParams* __restrict__ fp, unsigned long long* __restrict__ grid64) {
  extern __shared__ unsigned long long brick[];
  const int bin = xcd_remap(blockIdx.x, gridDim.x);
  if (bin >= b.nbins) return;
  const int s0 = starts[bin], s1 = starts[bin + 1];
  if (s0 == s1) return;
  const int bxi = bin % b.nbx, byi = (bin / b.nbx) % b.nby, bzi = bin / (b.nbx * b.nby);
  const int Lx = b.Lx, Ly = b.Ly, LxLy = b.Lx * b.Ly;
  const int vol = LxLy * b.Lz;
  for (int j = threadIdx.x; j < vol; j += blockDim.x) brick[j] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1;
  const int NZ = NZC > 0 ? NZC : 2 * g.nz + 1;
  const int NXY = NX * NY;
  const int KI = (NXY + 63) >> 6;
  const double S = fp->scale;
  const LaneRole role = lane_role(lane, NX, NY, NZ);

  int k_tx[kMaxK], k_ty[kMaxK], k_off[kMaxK];
  bool k_ok[kMaxK];
#pragma unroll
  for (int k = 0; k < kMaxK; k++) {
    const int idx = lane + 64 * k;
    k_ok[k] = (k < KI) && idx < NXY;
    k_tx[k] = k_ok[k] ? idx % NX : 0;
    k_ty[k] = k_ok[k] ? idx / NX : 0;
    k_off[k] = k_tx[k] + Lx * k_ty[k];
  }

  for (int i = s0 + wid; i < s1; i += nw) {
    const V4<CT> p = sorted[i];
    const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
    const int cx = (int)round(x * g.ihx) - g.nx;
    const int cy = (int)round(y * g.ihy) - g.ny;
    const int cz = (int)round(z * g.ihz) - g.nz;
    CT w = lane_weight<CT>(role, cx, cy, cz, x, y, z, g, inv2s2);
    if (normalize) {  // CpuGaussCharge.cpp:330-362, sums in j order
      CT sx = 0, sy = 0, sz = 0;
      for (int j = 0; j < NX; j++) sx += readlane(w, j);
      for (int j = 0; j < NY; j++) sy += readlane(w, NX + j);
      for (int j = 0; j < NZ; j++) sz += readlane(w, NX + NY + j);
      const CT nrm = role.axis == 0 ? (CT)pref1 / sx : role.axis == 1 ? (CT)pref1 / sy : (CT)pref1 / sz;
      w = w * nrm;
    }
    const int ox = wrapi(cx, g.X) - bxi * b.Bx;  // in [0, Bx)
    const int oy = wrapi(cy, g.Y) - byi * b.By;
    const int oz = wrapi(cz, g.Z) - bzi * b.Bz;
    const int base = ox + Lx * (oy + Ly * oz);
    const double qps = ((double)p.w * pref) * S;

    double wz[NZC > 0 ? NZC : 1];
    if (NZC > 0) {
#pragma unroll
      for (int t = 0; t < (NZC > 0 ? NZC : 1); t++) wz[t] = (double)readlane(w, NX + NY + t);
    }
#pragma unroll
    for (int k = 0; k < kMaxK; k++) {
      if (k >= KI) break;
      const double wxv = (double)shfl(w, k_tx[k]);
      const double wyv = (double)shfl(w, NX + k_ty[k]);
      const double qxy = (qps * wxv) * wyv;
      const int a0 = base + k_off[k];
      if (k_ok[k]) {
        if (NZC > 0) {
#pragma unroll
          for (int t = 0; t < (NZC > 0 ? NZC : 1); t++)
            atomicAdd(&brick[a0 + t * LxLy], (unsigned long long)fixed_round(qxy * wz[t]));
        } else {
          for (int t = 0; t < NZ; t++) {
            const double wzt = (double)readlane(w, NX + NY + t);
            atomicAdd(&brick[a0 + t * LxLy], (unsigned long long)fixed_round(qxy * wzt));
          }
        }
      }
    }
  }
  __syncthreads();

  const int X = g.X, Y = g.Y, Z = g.Z;
  const int gx0 = bxi * b.Bx, gy0 = byi * b.By, gz0 = bzi * b.Bz;
  for (int j = threadIdx.x; j < vol; j += blockDim.x) {
    const unsigned long long v = brick[j];
    if (v == 0ull) continue;
    const int lx = j % Lx;
    const int t = j / Lx;
    const int ly = t % Ly;
    const int lz = t / Ly;
    int gx = gx0 + lx, gy = gy0 + ly, gz = gz0 + lz;
    while (gx >= X) gx -= X;
    while (gy >= Y) gy -= Y;
    while (gz >= Z) gz -= Z;
    atomicAdd(&grid64[gx + (size_t)X * (gy + (size_t)Y * gz)], v);
  }
}

// Fallback when even a B=1 brick does not fit in LDS: the same particle
// loop adding straight into the global int64 grid.
template <typename CT>
__global__ void __launch_bounds__(256)
    k_spread_direct(Geom g, double inv2s2, double pref, int normalize, double pref1,
                    const int* __restrict__ starts, int nbins, const V4<CT>* __restrict__ sorted,
                    const FixedParams* __restrict__ fp, unsigned long long* __restrict__ grid64) {
  const int n = starts[nbins];
  const int lane = threadIdx.x & 63;
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (gridDim.x * blockDim.x) >> 6;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const double S = fp->scale;
  const LaneRole role = lane_role(lane, NX, NY, NZ);
  const bool fits = NX + NY + NZ <= 64;
  for (int i = gw; i < n; i += nwave) {
    const V4<CT> p = sorted[i];
    const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
    const int cx = (int)round(x * g.ihx) - g.nx;
    const int cy = (int)round(y * g.ihy) - g.ny;
    const int cz = (int)round(z * g.ihz) - g.nz;
    const double qps = ((double)p.w * pref) * S;
    CT w = fits ? lane_weight<CT>(role, cx, cy, cz, x, y, z, g, inv2s2) : (CT)0;
    CT sx = 1, sy = 1, sz = 1;
    if (normalize && fits) {
      sx = sy = sz = 0;
      for (int j = 0; j < NX; j++) sx += readlane(w, j);
      for (int j = 0; j < NY; j++) sy += readlane(w, NX + j);
      for (int j = 0; j < NZ; j++) sz += readlane(w, NX + NY + j);
    }
    const int wx0 = wrapi(cx, g.X), wy0 = wrapi(cy, g.Y), wz0 = wrapi(cz, g.Z);
    for (int idx = lane; idx < NX * NY * NZ; idx += 64) {
      const int tx = idx % NX, t2 = idx / NX, ty = t2 % NY, tz = t2 / NY;
      CT wx, wy, wzv;
      if (fits) {
        wx = shfl(w, tx);
        wy = shfl(w, NX + ty);
        wzv = shfl(w, NX + NY + tz);
      } else {  // very wide supports: evaluate the weights per point
        LaneRole r;
        r.axis = 0; r.j = tx; wx = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
        r.axis = 1; r.j = ty; wy = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
        r.axis = 2; r.j = tz; wzv = lane_weight<CT>(r, cx, cy, cz, x, y, z, g, inv2s2);
      }
      if (normalize) {
        wx = wx * ((CT)pref1 / sx);
        wy = wy * ((CT)pref1 / sy);
        wzv = wzv * ((CT)pref1 / sz);
      }
      const double v = ((qps * (double)wx) * (double)wy) * (double)wzv;
      const int gx = (wx0 + tx) % g.X, gy = (wy0 + ty) % g.Y, gz = (wz0 + tz) % g.Z;
      atomicAdd(&grid64[gx + (size_t)g.X * (gy + (size_t)g.Y * gz)],
                (unsigned long long)fixed_round(v));
    }
  }
}

template <typename T>
__global__ void k_fixed_to_grid(unsigned long long* __restrict__ g64, T* __restrict__ out,
                                size_t count, const FixedParams* __restrict__ fp, int clear) {
  const double inv = fp->inv_scale;
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

BinShape choose_bins(const Geom& g, size_t elem, size_t budget, int pref_edge) {
  BinShape b;
  int B[3] = {pref_edge, pref_edge, pref_edge};
  const int M[3] = {g.X, g.Y, g.Z};
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int a = 0; a < 3; a++)
    if (B[a] > M[a]) B[a] = M[a];
  // shrink the largest edge until the brick fits
  for (;;) {
    size_t bytes = (size_t)(B[0] + 2 * n[0]) * (B[1] + 2 * n[1]) * (B[2] + 2 * n[2]) * elem;
    if (bytes <= budget) break;
    int a = 0;
    for (int k = 1; k < 3; k++)
      if (B[k] > B[a]) a = k;
    if (B[a] == 1) {
      B[0] = B[1] = B[2] = 0;  // does not fit at all
      break;
    }
    B[a] = (B[a] + 1) / 2;
  }
  b.Bx = B[0];
  b.By = B[1];
  b.Bz = B[2];
  b.direct = 0;
  if (B[0] == 0) {  // one bin holding every particle
    b.Bx = g.X;
    b.By = g.Y;
    b.Bz = g.Z;
    b.nbx = b.nby = b.nbz = 1;
    b.nbins = 1;
    b.Lx = b.Ly = b.Lz = 0;
    b.direct = 1;
    return b;
  }
  b.nbx = (g.X + B[0] - 1) / B[0];
  b.nby = (g.Y + B[1] - 1) / B[1];
  b.nbz = (g.Z + B[2] - 1) / B[2];
  b.nbins = b.nbx * b.nby * b.nbz;
  b.Lx = B[0] + 2 * g.nx;
  b.Ly = B[1] + 2 * g.ny;
  b.Lz = B[2] + 2 * g.nz;
  return b;
}

hipError_t launch_bin(bool dbl, const Geom& g, const BinShape& b, int n, const void* xyzq,
                      Binning& w, int images, double pref, hipStream_t s) {
  k_bin_reset<<<grid_for(b.nbins + 1, 256), 256, 0, s>>>(w.counts, b.nbins + 1, w.fp);
  if (n > 0) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (dbl)
      k_bin_count<double><<<grid, 256, 0, s>>>(g, b, n, (const V4<double>*)xyzq, w.counts,
                                               w.atom_bin, w.atom_slot, w.fp);
    else
      k_bin_count<float><<<grid, 256, 0, s>>>(g, b, n, (const V4<float>*)xyzq, w.counts,
                                              w.atom_bin, w.atom_slot, w.fp);
  }
  k_bin_scan<<<1, 1024, 0, s>>>(w.counts, b.nbins, n, images, pref, w.fp);
  if (n > 0) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (dbl)
      k_bin_scatter<double><<<grid, 256, 0, s>>>(n, (const V4<double>*)xyzq, w.counts,
                                                 w.atom_bin, w.atom_slot, w.perm,
                                                 (V4<double>*)w.sorted);
    else
      k_bin_scatter<float><<<grid, 256, 0, s>>>(n, (const V4<float>*)xyzq, w.counts, w.atom_bin,
                                                w.atom_slot, w.perm, (V4<float>*)w.sorted);
  }
  return hipGetLastError();
}

template <typename CT>
static hipError_t spread_t(const Geom& g, const BinShape& b, double sigma, bool normalize,
                           Binning& w, unsigned long long* grid64, hipStream_t s) {
  const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
  const double pref = pow(2.0 * M_PI * sigma * sigma, -3.0 / 2.0);
  const double pref1 = sqrt(2.0 * M_PI * sigma * sigma);
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const bool lane_ok = NX + NY + NZ <= 64 && (NX * NY + 63) / 64 <= kMaxK;
  if (b.direct || !lane_ok) {
    k_spread_direct<CT><<<2048, 256, 0, s>>>(g, inv2s2, pref, normalize ? 1 : 0, pref1,
                                             w.counts, b.nbins, (const V4<CT>*)w.sorted, w.fp,
                                             grid64);
    return hipGetLastError();
  }
  const size_t lds = b.brick_bytes(8);
  const unsigned nwg = (unsigned)(((b.nbins + 7) / 8) * 8);
  const V4<CT>* srt = (const V4<CT>*)w.sorted;
  const int nz = NZ;
#define GSE_SPREAD_CASE(NZV)                                                                 \
  case NZV:                                                                                  \
    k_spread<CT, NZV><<<nwg, 256, lds, s>>>(g, b, inv2s2, pref, normalize ? 1 : 0, pref1,   \
                                             srt, w.counts, w.fp, grid64);                   \
    break;
  switch (nz) {
    GSE_SPREAD_CASE(7)
    GSE_SPREAD_CASE(9)
    GSE_SPREAD_CASE(11)
    GSE_SPREAD_CASE(13)
    GSE_SPREAD_CASE(15)
    GSE_SPREAD_CASE(17)
    default:
      k_spread<CT, 0><<<nwg, 256, lds, s>>>(g, b, inv2s2, pref, normalize ? 1 : 0, pref1, srt,
                                            w.counts, w.fp, grid64);
  }
#undef GSE_SPREAD_CASE
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
                                size_t count, const FixedParams* fp, bool clear, hipStream_t s) {
  const unsigned grid = grid_for(count, 256, 8192);
  if (dbl)
    k_fixed_to_grid<double><<<grid, 256, 0, s>>>((unsigned long long*)grid64, (double*)out,
                                                 count, fp, clear ? 1 : 0);
  else
    k_fixed_to_grid<float><<<grid, 256, 0, s>>>((unsigned long long*)grid64, (float*)out, count,
                                                fp, clear ? 1 : 0);
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

}  // namespace gse
