// Device-side helpers shared by the CDNA4 kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gse {

constexpr int kWave = 64;

// Fixed-point conversion of a double whose magnitude is below 2^51: adding
// 1.5*2^52 leaves round-to-nearest(v) in the low mantissa bits.
__device__ __forceinline__ long long fixed_round(double v) {
  const double magic = 6755399441055744.0;  // 1.5 * 2^52
  return (long long)__double_as_longlong(v + magic) -
         (long long)__double_as_longlong(magic);
}

// round-to-nearest(a*b) of the exact product (|a*b| below 2^51): one FMA
// adds the product to 1.5*2^52 with a single rounding.
__device__ __forceinline__ long long fixed_round_prod(double a, double b) {
  const double magic = 6755399441055744.0;  // 1.5 * 2^52
  return (long long)__double_as_longlong(__builtin_fma(a, b, magic)) -
         (long long)__double_as_longlong(magic);
}

// Uniform (wave-wide) broadcast of lane `src`'s value; src must be uniform.
__device__ __forceinline__ double readlane_d(double v, int src) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_f(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double readlane(double v, int s) { return readlane_d(v, s); }
__device__ __forceinline__ float readlane(float v, int s) { return readlane_f(v, s); }

// Workgroup schedule of the per-bin spread and gather.  xcd_remap gives XCD
// x the positions x*chunk .. x*chunk + chunk - 1, which its workgroups take
// in order; order[] holds, at those positions, the same bins sorted by
// descending charge count, so each XCD starts its longest bins first and
// the short ones fill the end of the launch.  Bins stay with their XCD (and
// its L2), and each bin's work, so every result, is the same in any order.
// One workgroup of BT >= 256 threads sorts eighth x (a counting sort on 256
// keys; counts of 255 and more share the first key, equal keys go in any
// order); sm: 2 * 256 + 4 unsigned of LDS.
constexpr int kOrderKeys = 256;
constexpr int kOrderWG = 8;  // one workgroup per XCD's eighth
__device__ __forceinline__ int order_key(int c) {
  return kOrderKeys - 1 - min(max(c, 0), kOrderKeys - 1);  // ascending = descending count
}
template <int BT>
__device__ void bin_order_eighth(const int* __restrict__ counts, int nbins, int x,
                                 int* __restrict__ order, unsigned* sm) {
  static_assert(BT >= kOrderKeys, "one thread per key in the scan");
  unsigned* hist = sm;
  unsigned* start = sm + kOrderKeys;
  unsigned* wtot = start + kOrderKeys;
  const int t = threadIdx.x;
  const int chunk = (((nbins + 7) >> 3) << 3) >> 3;  // xcd_remap's, for ((nbins + 7) / 8) * 8
  const int p0 = x * chunk, p1 = min(p0 + chunk, nbins);
  if (t < kOrderKeys) hist[t] = 0u;
  __syncthreads();
  for (int p = p0 + t; p < p1; p += BT) atomicAdd(&hist[order_key(counts[p])], 1u);
  __syncthreads();
  if (t < kOrderKeys) {  // exclusive scan of the histogram
    const unsigned h = hist[t];
    unsigned v = h;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned sh = __shfl_up(v, d, kWave);
      if ((t & 63) >= d) v += sh;
    }
    if ((t & 63) == 63) wtot[t >> 6] = v;
    start[t] = v - h;
  }
  __syncthreads();
  if (t < kOrderKeys) {
    unsigned add = 0u;
    for (int w = 0; w < (t >> 6); w++) add += wtot[w];
    start[t] += add;
  }
  __syncthreads();
  for (int p = p0 + t; p < p1; p += BT)
    order[p0 + (int)atomicAdd(&start[order_key(counts[p])], 1u)] = p;
}

// Lane shuffle with a per-lane source (ds_bpermute)
__device__ __forceinline__ double shfl_d(double v, int src) {
  int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
  int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float shfl_f(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double shfl(double v, int s) { return shfl_d(v, s); }
__device__ __forceinline__ float shfl(float v, int s) { return shfl_f(v, s); }

// Butterfly sum over the 64 lanes (every lane gets the total).  The order
// of additions is fixed, so the result is deterministic.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Wave total through DPP: xor-1 and xor-2 quad permutes, half-row and row
// mirrors (16-lane sums in every lane), then the four row sums read out and
// added in a fixed order.  The result is uniform and deterministic; no LDS
// traffic, unlike the ds_bpermute butterfly of wave_sum.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <typename T>
__device__ __forceinline__ T wave_total(T v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_move<0x141>(v);  // row_half_mirror: 8-lane sums
  v += dpp_move<0x140>(v);  // row_mirror: 16-lane sums
  const T r0 = readlane(v, 0), r1 = readlane(v, 16), r2 = readlane(v, 32), r3 = readlane(v, 48);
  return (r0 + r1) + (r2 + r3);
}

// Sum over each 16-lane row of the wave, in every lane of the row (fixed
// order, deterministic).
template <typename T>
__device__ __forceinline__ T row_total(T v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return v;
}

// Sum over each 32-lane half of the wave: lanes 0-31 get the sum of lanes
// 0-31, lanes 32-63 the sum of lanes 32-63 (fixed order, deterministic).
template <typename T>
__device__ __forceinline__ T half_total(T v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  const T r0 = readlane(v, 0), r1 = readlane(v, 16), r2 = readlane(v, 32), r3 = readlane(v, 48);
  return (threadIdx.x & 32) ? r2 + r3 : r0 + r1;
}

// Sums of four values over each 32-lane half of the wave in one transposed
// butterfly: after the lane ^ 1 and lane ^ 2 exchanges lane l holds value
// l & 3 summed over its quad, rotations by 4 and 8 lanes add the row's four
// quads, and one shuffle adds the half's other row.  Lane l returns the
// half's sum of value l & 3 (fixed order, deterministic).
__device__ __forceinline__ float group4_total(float a, float b, float c, float d, int lane) {
  const bool o1 = (lane & 1) != 0, o2 = (lane & 2) != 0;
  float p = o1 ? b : a;
  p += dpp_move<0xB1>(o1 ? a : b);
  float r = o1 ? d : c;
  r += dpp_move<0xB1>(o1 ? c : d);
  float v = o2 ? r : p;
  v += dpp_move<0x4E>(o2 ? p : r);
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x128>(v);  // row_ror:8
  return v + shfl(v, lane ^ 16);
}

// Non-negative modulo
__device__ __forceinline__ int wrapi(int i, int m) {
  int r = i % m;
  return r < 0 ? r + m : r;
}

// Explicitly non-contracted FP64 / FP32 arithmetic (the sphere test and cube
// selection must round exactly as the reference's separate operations).
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float sub_rn(float a, float b) { return __fsub_rn(a, b); }

__device__ __forceinline__ double gexp(double x) { return exp(x); }
__device__ __forceinline__ float gexp(float x) { return expf(x); }

// Error-free transformations for compensated sums (Ogita-Rump-Oishi):
// TwoSum (Knuth) and TwoProd via FMA.
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
  s = a + b;
  T bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(float a, float b, float& p, float& e) {
  p = a * b;
  e = __builtin_fmaf(a, b, -p);
}
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = __builtin_fma(a, b, -p);
}

// Compensated accumulator: value = s + c.
template <typename T>
struct KSum {
  T s, c;
  __device__ __forceinline__ KSum() : s(0), c(0) {}
  __device__ __forceinline__ void add(T v) {
    T t, e;
    two_sum(s, v, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void add_prod(T a, T b) {
    T p, ep, t, es;
    two_prod(a, b, p, ep);
    two_sum(s, p, t, es);
    s = t;
    c += ep + es;
  }
  __device__ __forceinline__ T value() const { return s + c; }
};

}  // namespace gse

namespace gse {

// Sum of partials[b*stride + c] over b < nblocks for c < NC, by one block in
// a fixed order (strided per-thread sums, butterfly, waves combined in index
// order).  Every thread receives the NC totals.  Needs blockDim.x <= 1024.
template <int NC>
__device__ __forceinline__ void reduce_partials_block(const double* __restrict__ partials,
                                                      int nblocks, int stride, double (&tot)[NC]) {
  __shared__ double sh_red[NC][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[b * stride + c];
    s = wave_sum(s);
    if (lane == 0) sh_red[c][wid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; c++) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh_red[c][w];
    tot[c] = s;
  }
}

// Energy / k-space virial from the k-pass partials ([nblocks][ncomp], ncomp 1
// or 7): energy[0] = pref*E, virial[3a+b] = -delta_ab E - pref*K_ab.
__device__ __forceinline__ void kpass_finalize_block(const double* __restrict__ partials,
                                                     int nblocks, int ncomp, double pref,
                                                     double* energy, double* virial) {
  const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  if (ncomp == 7) {
    double t[7];
    reduce_partials_block<7>(partials, nblocks, 7, t);
    const double E = t[0] * pref;
    if (threadIdx.x == 0 && energy) energy[0] = E;
    if (virial && threadIdx.x < 9) {
      const int i = threadIdx.x;
      virial[i] = -t[1 + map[i]] * pref - ((i == 0 || i == 4 || i == 8) ? E : 0.0);
    }
  } else {
    double t[1];
    reduce_partials_block<1>(partials, nblocks, ncomp, t);
    if (threadIdx.x == 0 && energy) energy[0] = t[0] * pref;
  }
}

}  // namespace gse
