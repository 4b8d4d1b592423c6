// Comparators beside the GSE path, on the same device infrastructure:
//
// SPME reciprocal space (the reference's cuda/CudaPMERecip.cu, order 4/6/8):
//   k_pme_qmax      max |q| and the rejected-particle count of one call
//   k_pme_spread    B-spline charge spread into the int64 fixed-point grid
//                   (one lane per (x, y) spline offset, a loop over z; the
//                   same exponent rule as the GSE spread, so the grid is
//                   bit-identical for any particle order); the conversion to
//                   the compute type is the GSE path's k_fixed_to_grid
//   k_pme_scalar_sum  the spectrum times the SPME kernel with FP64 energy and
//                   virial partial sums per workgroup (fixed order)
//   k_pme_finalize  energy and the 3x3 virial from the partial sums
//   k_pme_gather    energy gradient per particle: one lane group per particle
//                   (16 lanes for order 4, a wave for 6 and 8), lanes on the
//                   (x, y) offsets, a loop over z, a fixed-order butterfly
// The spline position, weights and moduli follow CudaPMERecip.cu:266-327,
// :436-588, :946-1256 and :1612-1890 term for term.
//
// Ewald direct space (CpuEwaldDirect.cpp:45-131):
//   k_ew_cell_count / k_ew_scan / k_ew_scatter / k_ew_sort
//                   particles into cells of edge >= rcut, each cell's list in
//                   ascending particle index (deterministic order)
//   k_ew_force      one workgroup per cell: four lanes share one particle
//                   (each a quarter of its neighbours, summed across the four
//                   in a fixed order), neighbour cells are staged through
//                   LDS, the pair term of
//                   the reference is evaluated with its operand order and
//                   minimum-image loops; FP64 energy partial per cell
#include <math.h>

#include "gse_device.h"
#include "gse_internal.h"

namespace gse {

template <typename T>
struct alignas(4 * sizeof(T)) P4 {
  T x, y, z, w;
};

__global__ void k_clear_stats(BinStats* st) {
  st->qmax_bits = 0ull;
  st->invalid = 0;
  st->overflow = 0;
  st->spill = 0;
}

__global__ void k_zero_i32(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------
// SPME
// ---------------------------------------------------------------------------
// fractional grid position (CudaPMERecip.cu:471-476, in FP64)
__device__ __forceinline__ void pme_pos(double x, double recip, int nfft, int& i0,
                                        double& frac) {
  const double w = x * recip + 2.0;
  const double fr = (double)nfft * (w - (floor(w + 0.5) - 0.5));
  const int fi = (int)fr;
  frac = fr - (double)fi;
  i0 = fi;
}

// B-spline weights and derivatives (CudaPMERecip.cu:266-327)
template <int ORDER>
__device__ __forceinline__ void pme_theta(double w, double* th, double* dth) {
  th[ORDER - 1] = 0.0;
  th[1] = w;
  th[0] = 1.0 - w;
#pragma unroll
  for (int k = 3; k <= ORDER - 1; k++) {
    const double div = 1.0 / (double)(k - 1);
    th[k - 1] = div * w * th[k - 2];
#pragma unroll
    for (int j = 1; j <= k - 2; j++)
      th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 2; j <= ORDER; j++) dth[j - 1] = th[j - 2] - th[j - 1];
  const double div = 1.0 / (double)(ORDER - 1);
  th[ORDER - 1] = div * w * th[ORDER - 2];
#pragma unroll
  for (int j = 1; j <= ORDER - 2; j++)
    th[ORDER - j - 1] = div * ((w + j) * th[ORDER - j - 2] + (ORDER - j - w) * th[ORDER - j - 1]);
  th[0] = div * (1.0 - w) * th[0];
}

// A particle takes part when its record is finite and its scaled coordinates
// are small enough for the integer grid position.
template <typename CT>
__device__ __forceinline__ bool pme_valid(const P4<CT>& r, double rx, double ry, double rz) {
  const double x = r.x, y = r.y, z = r.z, q = r.w;
  return isfinite(x) && isfinite(y) && isfinite(z) && isfinite(q) && fabs(x * rx) < 1e9 &&
         fabs(y * ry) < 1e9 && fabs(z * rz) < 1e9;
}

template <typename CT>
__global__ __launch_bounds__(256) void k_pme_qmax(const P4<CT>* __restrict__ xyzq, int n,
                                                  double rx, double ry, double rz,
                                                  BinStats* __restrict__ st) {
  double qm = 0.0;
  int bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const P4<CT> r = xyzq[i];
    if (pme_valid(r, rx, ry, rz))
      qm = fmax(qm, fabs((double)r.w));
    else
      bad++;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    qm = fmax(qm, __shfl_xor(qm, m, kWave));
    bad += __shfl_xor(bad, m, kWave);
  }
  if ((threadIdx.x & 63) == 0) {
    if (qm > 0.0) atomicMax(&st->qmax_bits, (unsigned long long)__double_as_longlong(qm));
    if (bad) atomicAdd(&st->invalid, bad);
  }
}

// lanes per particle: order^2 (x, y) offsets rounded up to 16 or 64
template <int ORDER>
struct PmeLanes {
  static constexpr int G = ORDER * ORDER <= 16 ? 16 : 64;
};

template <int ORDER, typename CT>
__global__ __launch_bounds__(256) void k_pme_spread(const P4<CT>* __restrict__ xyzq, int n,
                                                    int X, int Y, int Z, double rx, double ry,
                                                    double rz, const BinStats* __restrict__ st,
                                                    unsigned long long* __restrict__ g64) {
  constexpr int G = PmeLanes<ORDER>::G;
  constexpr int PER = 256 / G;
  const int lane = threadIdx.x & (G - 1);
  const int a = lane % ORDER, b = lane / ORDER;
  const BinStats s = *st;
  const double scale = ldexp(1.0, fixed_exponent(s.qmax_bits, n - s.invalid, 1, 1.0));
  for (int p = blockIdx.x * PER + (int)(threadIdx.x / G); p < n; p += gridDim.x * PER) {
    const P4<CT> r = xyzq[p];
    if (!pme_valid(r, rx, ry, rz) || lane >= ORDER * ORDER) continue;
    int ix, iy, iz;
    double fx, fy, fz;
    pme_pos((double)r.x, rx, X, ix, fx);
    pme_pos((double)r.y, ry, Y, iy, fy);
    pme_pos((double)r.z, rz, Z, iz, fz);
    double tx[ORDER], ty[ORDER], tz[ORDER], d[ORDER];
    pme_theta<ORDER>(fx, tx, d);
    pme_theta<ORDER>(fy, ty, d);
    pme_theta<ORDER>(fz, tz, d);
    // ((q tx) ty) tz as the reference orders it; the exact power-of-two
    // scale changes no rounding
    const double qxy = (double)r.w * scale * tx[a] * ty[b];
    int xi = ix + a, yi = iy + b;
    if (xi >= X) xi -= X;
    if (yi >= Y) yi -= Y;
    const size_t row = (size_t)xi + (size_t)yi * X;
#pragma unroll
    for (int c = 0; c < ORDER; c++) {
      int zi = iz + c;
      if (zi >= Z) zi -= Z;
      const long long v = fixed_round_prod(qxy, tz[c]);
      atomicAdd(&g64[row + (size_t)zi * X * Y], (unsigned long long)v);
    }
  }
}

struct PmeSum {
  int X, Y, Z;
  double rx, ry, rz, fac, piv_inv;
};

__device__ __forceinline__ double pme_eterm(const PmeSum& a, int k1, int k2, int k3,
                                            const double* p1, const double* p2,
                                            const double* p3, double* mh) {
  const int nf1 = a.X / 2 + a.X % 2, nf2 = a.Y / 2 + a.Y % 2, nf3 = a.Z / 2 + a.Z % 2;
  const int m1 = k1 >= nf1 ? k1 - a.X : k1, m2 = k2 >= nf2 ? k2 - a.Y : k2,
            m3 = k3 >= nf3 ? k3 - a.Z : k3;
  mh[0] = a.rx * m1;
  mh[1] = a.ry * m2;
  mh[2] = a.rz * m3;
  const double msq = mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2];
  return exp(-a.fac * msq) * a.piv_inv * p1[k1] * p2[k2] * p3[k3] / msq;
}

// CudaPMERecip.cu:946-1256: the element and, for x index >= 1, its
// conjugate partner enter the energy and virial (the x Nyquist plane of an
// even grid twice, as in the reference).  Partial sums per workgroup in
// partials[blockIdx][7] (energy, then the six upper-triangle moments).
template <typename CT>
__global__ __launch_bounds__(256) void k_pme_scalar_sum(PmeSum a, const double* __restrict__ p1,
                                                        const double* __restrict__ p2,
                                                        const double* __restrict__ p3,
                                                        CT* __restrict__ spec, int want,
                                                        double* __restrict__ partials) {
  const int size1 = a.X / 2 + 1;
  const long long K = (long long)size1 * a.Y * a.Z;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < K;
       e += (long long)gridDim.x * blockDim.x) {
    CT* q = spec + 2 * e;
    if (e == 0) {
      q[0] = (CT)0;
      q[1] = (CT)0;
      continue;
    }
    const int k1 = (int)(e % size1);
    const long long r = e / size1;
    const int k2 = (int)(r % a.Y), k3 = (int)(r / a.Y);
    double mh[3];
    const double et = pme_eterm(a, k1, k2, k3, p1, p2, p3, mh);
    if (want) {
      const double q2 = (double)q[0] * (double)q[0] + (double)q[1] * (double)q[1];
      for (int part = 0; part < (k1 >= 1 ? 2 : 1); part++) {
        double ee = et, m[3] = {mh[0], mh[1], mh[2]};
        if (part == 1)
          ee = pme_eterm(a, a.X - k1, (a.Y - k2) % a.Y, (a.Z - k3) % a.Z, p1, p2, p3, m);
        const double msq = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        const double t1 = ee * q2, vt = 2.0 * (a.fac + 1.0 / msq), t2 = t1 * vt;
        acc[0] += t1;
        acc[1] += t1 * (vt * m[0] * m[0] - 1.0);
        acc[2] += t2 * m[0] * m[1];
        acc[3] += t2 * m[0] * m[2];
        acc[4] += t1 * (vt * m[1] * m[1] - 1.0);
        acc[5] += t2 * m[1] * m[2];
        acc[6] += t1 * (vt * m[2] * m[2] - 1.0);
      }
    }
    q[0] = (CT)((double)q[0] * et);
    q[1] = (CT)((double)q[1] * et);
  }
  if (!want) return;
  __shared__ double sh[4][7];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 7; c++) {
    const double t = wave_sum(acc[c]);
    if ((threadIdx.x & 63) == 0) sh[w][c] = t;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const int c = threadIdx.x;
    partials[blockIdx.x * 7 + c] = (sh[0][c] + sh[1][c]) + (sh[2][c] + sh[3][c]);
  }
}

// E = ccelec/2 sum, virial_ab = -ccelec/2 sum (CudaPMERecip.cu:1222-1239)
// the workgroup's threads sum strided partials, then the waves' sums are
// added in wave order (fixed order, deterministic)
__global__ void __launch_bounds__(256)
    k_pme_finalize(const double* __restrict__ partials, int nblocks, double ccelec,
                   double* energy, double* virial) {
  double sh[7];
  reduce_partials_block<7>(partials, nblocks, 7, sh);
  if (threadIdx.x == 0) {
    const double h = 0.5 * ccelec;
    if (energy) energy[0] = h * sh[0];
    if (virial) {
      virial[0] = -h * sh[1];
      virial[1] = -h * sh[2];
      virial[2] = -h * sh[3];
      virial[3] = -h * sh[2];
      virial[4] = -h * sh[4];
      virial[5] = -h * sh[5];
      virial[6] = -h * sh[3];
      virial[7] = -h * sh[5];
      virial[8] = -h * sh[6];
    }
  }
}

// CudaPMERecip.cu:1612-1890: dE/dr_a = q ccelec (M_a/L_a) sum dtheta_a theta
// theta Qc; written times `sign` into out{x,y,z}[i].
template <int ORDER, typename CT, typename AT>
__global__ __launch_bounds__(256) void k_pme_gather(const P4<CT>* __restrict__ xyzq, int n,
                                                    int X, int Y, int Z, double rx, double ry,
                                                    double rz, double ccelec, double sign,
                                                    const CT* __restrict__ Qc, AT* __restrict__ ox,
                                                    AT* __restrict__ oy, AT* __restrict__ oz) {
  constexpr int G = PmeLanes<ORDER>::G;
  constexpr int PER = 256 / G;
  const int lane = threadIdx.x & (G - 1);
  const int a = lane % ORDER, b = lane / ORDER;
  const bool live = lane < ORDER * ORDER;
  // uniform trip count over the group so the butterfly sees every lane
  const int pstart = blockIdx.x * PER + (int)(threadIdx.x / G);
  const int nloop = (n + gridDim.x * PER - 1) / (gridDim.x * PER);
  for (int it = 0; it < nloop; it++) {
    const int p = pstart + it * gridDim.x * PER;
    CT f1 = 0, f2 = 0, f3 = 0;
    bool ok = false;
    P4<CT> r{};
    if (p < n) {
      r = xyzq[p];
      ok = pme_valid(r, rx, ry, rz);
    }
    if (ok && live) {
      int ix, iy, iz;
      double fx, fy, fz;
      pme_pos((double)r.x, rx, X, ix, fx);
      pme_pos((double)r.y, ry, Y, iy, fy);
      pme_pos((double)r.z, rz, Z, iz, fz);
      double tx[ORDER], ty[ORDER], tz[ORDER], dx[ORDER], dy[ORDER], dz[ORDER];
      pme_theta<ORDER>(fx, tx, dx);
      pme_theta<ORDER>(fy, ty, dy);
      pme_theta<ORDER>(fz, tz, dz);
      int xi = ix + a, yi = iy + b;
      if (xi >= X) xi -= X;
      if (yi >= Y) yi -= Y;
      const CT* row = Qc + (size_t)xi + (size_t)yi * X;
      const CT w1 = (CT)(dx[a] * ty[b]), w2 = (CT)(tx[a] * dy[b]), w3 = (CT)(tx[a] * ty[b]);
#pragma unroll
      for (int c = 0; c < ORDER; c++) {
        int zi = iz + c;
        if (zi >= Z) zi -= Z;
        const CT g = row[(size_t)zi * X * Y];
        const CT zc = (CT)tz[c], zd = (CT)dz[c];
        f1 += w1 * zc * g;
        f2 += w2 * zc * g;
        f3 += w3 * zd * g;
      }
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
      f1 += __shfl_xor(f1, m, kWave);
      f2 += __shfl_xor(f2, m, kWave);
      f3 += __shfl_xor(f3, m, kWave);
    }
    if (lane == 0 && p < n) {
      const double q = ok ? (double)r.w * ccelec * sign : 0.0;
      ox[p] = (AT)(q * rx * (double)f1 * X);
      oy[p] = (AT)(q * ry * (double)f2 * Y);
      oz[p] = (AT)(q * rz * (double)f3 * Z);
    }
  }
}

static unsigned blocks_for(long long count, unsigned threads, unsigned cap) {
  long long b = (count + threads - 1) / threads;
  if (b < 1) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

hipError_t launch_pme_qmax(bool dbl, const void* xyzq, int n, const double* recip3,
                           BinStats* stats, hipStream_t s) {
  k_clear_stats<<<1, 1, 0, s>>>(stats);
  if (n <= 0) return hipGetLastError();
  const unsigned grid = blocks_for(n, 256, 1024);
  if (dbl)
    k_pme_qmax<double><<<grid, 256, 0, s>>>((const P4<double>*)xyzq, n, recip3[0], recip3[1],
                                            recip3[2], stats);
  else
    k_pme_qmax<float><<<grid, 256, 0, s>>>((const P4<float>*)xyzq, n, recip3[0], recip3[1],
                                           recip3[2], stats);
  return hipGetLastError();
}

template <int ORDER>
static void pme_spread_o(bool dbl, const void* xyzq, int n, int X, int Y, int Z,
                         const double* r, const BinStats* st, unsigned long long* g64,
                         hipStream_t s) {
  constexpr int PER = 256 / PmeLanes<ORDER>::G;
  const unsigned grid = blocks_for((n + PER - 1) / PER, 1, 8192);
  if (dbl)
    k_pme_spread<ORDER, double><<<grid, 256, 0, s>>>((const P4<double>*)xyzq, n, X, Y, Z, r[0],
                                                     r[1], r[2], st, g64);
  else
    k_pme_spread<ORDER, float><<<grid, 256, 0, s>>>((const P4<float>*)xyzq, n, X, Y, Z, r[0],
                                                    r[1], r[2], st, g64);
}

hipError_t launch_pme_spread(bool dbl, int order, const void* xyzq, int n, int X, int Y, int Z,
                             const double* recip3, const BinStats* stats,
                             unsigned long long* g64, hipStream_t s) {
  if (n <= 0) return hipGetLastError();
  switch (order) {
    case 4: pme_spread_o<4>(dbl, xyzq, n, X, Y, Z, recip3, stats, g64, s); break;
    case 6: pme_spread_o<6>(dbl, xyzq, n, X, Y, Z, recip3, stats, g64, s); break;
    case 8: pme_spread_o<8>(dbl, xyzq, n, X, Y, Z, recip3, stats, g64, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pme_scalar_sum(bool dbl, int X, int Y, int Z, const double* recip3,
                                 double kappa, const double* prefac, void* spec, bool want,
                                 double* partials, hipStream_t s) {
  PmeSum a;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.rx = recip3[0];
  a.ry = recip3[1];
  a.rz = recip3[2];
  a.fac = M_PI * M_PI / (kappa * kappa);
  a.piv_inv = recip3[0] * recip3[1] * recip3[2] / M_PI;
  const double* p1 = prefac;
  const double* p2 = prefac + X;
  const double* p3 = prefac + X + Y;
  if (dbl)
    k_pme_scalar_sum<double><<<kPmeSumBlocks, 256, 0, s>>>(a, p1, p2, p3, (double*)spec,
                                                           want ? 1 : 0, partials);
  else
    k_pme_scalar_sum<float><<<kPmeSumBlocks, 256, 0, s>>>(a, p1, p2, p3, (float*)spec,
                                                          want ? 1 : 0, partials);
  return hipGetLastError();
}

hipError_t launch_pme_finalize(const double* partials, double ccelec, double* energy,
                               double* virial, hipStream_t s) {
  k_pme_finalize<<<1, 256, 0, s>>>(partials, kPmeSumBlocks, ccelec, energy, virial);
  return hipGetLastError();
}

template <int ORDER, typename CT>
static void pme_gather_o(bool at_dbl, const void* xyzq, int n, int X, int Y, int Z,
                         const double* r, double ccelec, double sign, const void* Qc, void* ox,
                         void* oy, void* oz, hipStream_t s) {
  constexpr int PER = 256 / PmeLanes<ORDER>::G;
  const unsigned grid = blocks_for((n + PER - 1) / PER, 1, 8192);
  if (at_dbl)
    k_pme_gather<ORDER, CT, double><<<grid, 256, 0, s>>>(
        (const P4<CT>*)xyzq, n, X, Y, Z, r[0], r[1], r[2], ccelec, sign, (const CT*)Qc,
        (double*)ox, (double*)oy, (double*)oz);
  else
    k_pme_gather<ORDER, CT, float><<<grid, 256, 0, s>>>(
        (const P4<CT>*)xyzq, n, X, Y, Z, r[0], r[1], r[2], ccelec, sign, (const CT*)Qc,
        (float*)ox, (float*)oy, (float*)oz);
}

hipError_t launch_pme_gather(bool ct_dbl, bool at_dbl, int order, const void* xyzq, int n, int X,
                             int Y, int Z, const double* recip3, double ccelec, double sign,
                             const void* Qc, void* ox, void* oy, void* oz, hipStream_t s) {
  if (n <= 0) return hipGetLastError();
#define GSE_PME_GATHER(O)                                                                     \
  if (ct_dbl)                                                                                 \
    pme_gather_o<O, double>(at_dbl, xyzq, n, X, Y, Z, recip3, ccelec, sign, Qc, ox, oy, oz, s); \
  else                                                                                        \
    pme_gather_o<O, float>(at_dbl, xyzq, n, X, Y, Z, recip3, ccelec, sign, Qc, ox, oy, oz, s);
  switch (order) {
    case 4: GSE_PME_GATHER(4) break;
    case 6: GSE_PME_GATHER(6) break;
    case 8: GSE_PME_GATHER(8) break;
    default: return hipErrorInvalidValue;
  }
#undef GSE_PME_GATHER
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ewald direct space
// ---------------------------------------------------------------------------
// cell of one coordinate: wrapped into [0, L), cells of edge L/nc
__device__ __forceinline__ int ew_cell1(double x, double L, int nc) {
  double w = x - L * floor(x / L);
  if (!(w >= 0.0 && w < L)) return 0;
  const int c = (int)(w * ((double)nc / L));
  return c < nc ? c : nc - 1;
}

__global__ void k_ew_cell_count(const P4<double>* __restrict__ xyzq, int n, EwCells c,
                                int* __restrict__ cell_of, int* __restrict__ counts) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const P4<double> r = xyzq[i];
    const int cx = ew_cell1(r.x, c.box[0], c.nc[0]), cy = ew_cell1(r.y, c.box[1], c.nc[1]),
              cz = ew_cell1(r.z, c.box[2], c.nc[2]);
    const int cell = cx + (cy + cz * c.nc[1]) * c.nc[0];
    cell_of[i] = cell;
    atomicAdd(&counts[cell], 1);
  }
}

// exclusive scan of the cell counts by one workgroup; fill[] cleared
__global__ __launch_bounds__(1024) void k_ew_scan(const int* __restrict__ counts, int ncell,
                                                  int* __restrict__ starts,
                                                  int* __restrict__ fill) {
  __shared__ int sh[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < ncell; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < ncell ? counts[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < ncell) {
      starts[i] = carry + sh[threadIdx.x] - v;
      fill[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
}

__global__ void k_ew_scatter(int n, const int* __restrict__ cell_of,
                             const int* __restrict__ starts, int* __restrict__ fill,
                             int* __restrict__ tmp) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = cell_of[i];
    tmp[starts[c] + atomicAdd(&fill[c], 1)] = i;
  }
}

// each cell's particles in ascending index: the rank of a slot is the number
// of the cell's particles with a smaller index
__global__ void k_ew_sort(int n, const int* __restrict__ cell_of, const int* __restrict__ starts,
                          const int* __restrict__ counts, const int* __restrict__ tmp,
                          const P4<double>* __restrict__ xyzq, int* __restrict__ sidx,
                          P4<double>* __restrict__ sx) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int i = tmp[s];
    const int c = cell_of[i];
    const int st = starts[c], cnt = counts[c];
    int rank = 0;
    for (int k = 0; k < cnt; k++) rank += tmp[st + k] < i;
    sidx[st + rank] = i;
    sx[st + rank] = xyzq[i];
  }
}

// the reference's minimum-image loops (CpuEwaldDirect.cpp:93-98), bounded:
// after four steps a coordinate still outside is folded in one go (only
// coordinates many boxes away, or non-finite ones, get there)
__device__ __forceinline__ double ew_image(double d, double L, double h) {
  for (int it = 0; it < 4 && d < -h; it++) d += L;
  for (int it = 0; it < 4 && d > h; it++) d -= L;
  if (d < -h || d > h) d -= L * rint(d / L);
  return d;
}

constexpr int kEwBlock = 128;  // particles of a cell per pass
constexpr int kEwSplit = 4;    // threads per particle, each a quarter of the neighbours
constexpr int kEwThreads = kEwBlock * kEwSplit;

// One workgroup per cell; kEwSplit consecutive lanes share one particle of
// the cell and take every kEwSplit-th neighbour of each staged neighbour
// cell, and their sums are added across the four lanes in a fixed order.
__global__ __launch_bounds__(kEwThreads) void k_ew_force(
    EwCells c, EwPair pr, const int* __restrict__ starts, const int* __restrict__ counts,
    const int* __restrict__ sidx, const P4<double>* __restrict__ sx,
    const int* __restrict__ xofs, const int* __restrict__ xlist, double* __restrict__ gx,
    double* __restrict__ gy, double* __restrict__ gz, double* __restrict__ partials) {
  __shared__ P4<double> shp[kEwBlock];
  __shared__ int shi[kEwBlock];
  __shared__ double she[kEwThreads / 64];
  const int cell = blockIdx.x;
  const int cx = cell % c.nc[0], cy = (cell / c.nc[0]) % c.nc[1], cz = cell / (c.nc[0] * c.nc[1]);
  const int st = starts[cell], cnt = counts[cell];
  const int lo0 = c.nc[0] >= 3 ? -1 : 0, lo1 = c.nc[1] >= 3 ? -1 : 0, lo2 = c.nc[2] >= 3 ? -1 : 0;
  const int slot = threadIdx.x / kEwSplit, sp = threadIdx.x % kEwSplit;
  double esum = 0.0;
  for (int base = 0; base < cnt; base += kEwBlock) {
    const int t = base + slot;
    const bool own = t < cnt;
    const int a = own ? sidx[st + t] : -1;
    P4<double> pa{0.0, 0.0, 0.0, 0.0};
    int x0 = 0, x1 = 0;
    if (own) {
      pa = sx[st + t];
      if (xofs) {
        x0 = xofs[a];
        x1 = xofs[a + 1];
      }
    }
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, e = 0.0;
    for (int dz = lo2; dz <= -lo2; dz++)
      for (int dy = lo1; dy <= -lo1; dy++)
        for (int dx = lo0; dx <= -lo0; dx++) {
          const int nx = (cx + dx + c.nc[0]) % c.nc[0], ny = (cy + dy + c.nc[1]) % c.nc[1],
                    nz = (cz + dz + c.nc[2]) % c.nc[2];
          const int nb = nx + (ny + nz * c.nc[1]) * c.nc[0];
          const int nst = starts[nb], ncnt = counts[nb];
          for (int kb = 0; kb < ncnt; kb += kEwBlock) {
            __syncthreads();
            if ((int)threadIdx.x < kEwBlock && kb + (int)threadIdx.x < ncnt) {
              shp[threadIdx.x] = sx[nst + kb + threadIdx.x];
              shi[threadIdx.x] = sidx[nst + kb + threadIdx.x];
            }
            __syncthreads();
            const int m = min(kEwBlock, ncnt - kb);
            if (!own) continue;
            for (int k = sp; k < m; k += kEwSplit) {
              const int b = shi[k];
              if (b == a) continue;
              const P4<double> pb = shp[k];
              const double ddx = ew_image(pa.x - pb.x, c.box[0], 0.5 * c.box[0]);
              const double ddy = ew_image(pa.y - pb.y, c.box[1], 0.5 * c.box[1]);
              const double ddz = ew_image(pa.z - pb.z, c.box[2], 0.5 * c.box[2]);
              const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
              if (!(r2 < pr.rcut2)) continue;
              bool excl = false;
              for (int q = x0; q < x1; q++) excl = excl || xlist[q] == b;
              if (excl) continue;
              // qq = (q_j ccelec) q_i with j the smaller index (:69, :101)
              const double qq = a < b ? (pa.w * pr.ccelec) * pb.w : (pb.w * pr.ccelec) * pa.w;
              const double r = sqrt(r2);
              const double erfcval = erfc(r * pr.inv_sqrt2sigma);
              const double expval = exp(-r2 * pr.inv_2sigmasq);
              const double f = -qq * (pr.two_sqrtpi * pr.inv_sqrt2sigma * expval + erfcval / r) / r2;
              g0 += f * ddx;
              g1 += f * ddy;
              g2 += f * ddz;
              e += qq * erfcval / r;
            }
          }
        }
    // the particle's four partial sums, added in a fixed order (xor 1, then
    // xor 2: every lane of the four ends with the same value)
#pragma unroll
    for (int m = 1; m < kEwSplit; m <<= 1) {
      g0 += __shfl_xor(g0, m, kWave);
      g1 += __shfl_xor(g1, m, kWave);
      g2 += __shfl_xor(g2, m, kWave);
    }
    if (own && sp == 0) {
      gx[a] = g0;
      gy[a] = g1;
      gz[a] = g2;
    }
    esum += e;
  }
  // every pair was met from both ends: half of the sum per cell
  const double w = wave_sum(esum);
  if ((threadIdx.x & 63) == 0) she[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int k = 0; k < kEwThreads / 64; k++) tot += she[k];
    partials[cell] = 0.5 * tot;
  }
}

// energy = sum of the cell partials (cell order) + the self term
__global__ void k_ew_finalize(const double* __restrict__ partials, int ncell,
                              const double* __restrict__ self2, double* energy) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < ncell; i += 256) s += partials[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) energy[0] = sh[0] + self2[0];
}

hipError_t launch_ew_cells(const void* xyzq, int n, const EwCells& c, int* cell_of, int* counts,
                           int* starts, int* fill, int* tmp, int* sidx, void* sx,
                           hipStream_t s) {
  k_zero_i32<<<blocks_for(c.ncell, 256, 1024), 256, 0, s>>>(counts, c.ncell);
  if (n > 0) {
    const unsigned g = blocks_for(n, 256, 4096);
    k_ew_cell_count<<<g, 256, 0, s>>>((const P4<double>*)xyzq, n, c, cell_of, counts);
  }
  k_ew_scan<<<1, 1024, 0, s>>>(counts, c.ncell, starts, fill);
  if (n > 0) {
    const unsigned g = blocks_for(n, 256, 4096);
    k_ew_scatter<<<g, 256, 0, s>>>(n, cell_of, starts, fill, tmp);
    k_ew_sort<<<g, 256, 0, s>>>(n, cell_of, starts, counts, tmp, (const P4<double>*)xyzq, sidx,
                                (P4<double>*)sx);
  }
  return hipGetLastError();
}

hipError_t launch_ew_force(const EwCells& c, const EwPair& p, const int* starts,
                           const int* counts, const int* sidx, const void* sx, const int* xofs,
                           const int* xlist, double* gx, double* gy, double* gz,
                           double* partials, hipStream_t s) {
  k_ew_force<<<c.ncell, kEwThreads, 0, s>>>(c, p, starts, counts, sidx, (const P4<double>*)sx,
                                            xofs, xlist, gx, gy, gz, partials);
  return hipGetLastError();
}

hipError_t launch_ew_finalize(const double* partials, int ncell, const double* self2,
                              double* energy, hipStream_t s) {
  k_ew_finalize<<<1, 256, 0, s>>>(partials, ncell, self2, energy);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ewald reciprocal sum (CpuEwaldRecip.cpp:87-147), FP64
// ---------------------------------------------------------------------------
// k-vectors of the cube |kx|,|ky|,|kz| <= kcut, x fastest as the reference
// loops, kept when 0 < kx^2+ky^2+kz^2 <= kcut^2.  One thread per k: its
// structure factor S(k) = sum q (cos, sin)(k.r) over all particles (staged
// through LDS, particle order), pek = pref exp(-k^2 sigma^2/2)/k^2, and the
// energy term pek |S|^2 (partial sums per workgroup).
constexpr int kErBlock = 256;

__device__ __forceinline__ bool er_kvec(const ErArgs& a, int idx, double* kv) {
  const int n1 = 2 * a.kcut + 1;
  const int kx = idx % n1 - a.kcut, ky = (idx / n1) % n1 - a.kcut, kz = idx / (n1 * n1) - a.kcut;
  const int k2 = kx * kx + ky * ky + kz * kz;
  if ((kx == 0 && ky == 0 && kz == 0) || k2 > a.kcut * a.kcut) return false;
  kv[0] = 2.0 * M_PI * ((double)kx) * a.ib[0];
  kv[1] = 2.0 * M_PI * ((double)ky) * a.ib[1];
  kv[2] = 2.0 * M_PI * ((double)kz) * a.ib[2];
  return true;
}

__global__ __launch_bounds__(kErBlock) void k_er_sfac(ErArgs a, const P4<double>* __restrict__ xyzq,
                                                      int n, double* __restrict__ sk,
                                                      double* __restrict__ partials) {
  __shared__ P4<double> sh[kErBlock];
  __shared__ double she[kErBlock / 64];
  const int idx = blockIdx.x * kErBlock + threadIdx.x;
  double kv[3] = {0, 0, 0};
  const bool live = idx < a.ncube && er_kvec(a, idx, kv);
  double re = 0.0, im = 0.0;
  for (int base = 0; base < n; base += kErBlock) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) sh[threadIdx.x] = xyzq[base + threadIdx.x];
    __syncthreads();
    const int m = min(kErBlock, n - base);
    if (live)
      for (int i = 0; i < m; i++) {
        const P4<double> p = sh[i];
        const double kr = kv[0] * p.x + kv[1] * p.y + kv[2] * p.z;
        double sn, cs;
        sincos(kr, &sn, &cs);
        re += p.w * cs;
        im += p.w * sn;
      }
  }
  double e = 0.0;
  if (live) {
    const double ksq = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    const double pek = a.pref * exp(-ksq * a.half_sigmasq) / ksq;
    e = pek * (re * re + im * im);
    sk[3 * (size_t)idx + 0] = re;
    sk[3 * (size_t)idx + 1] = im;
    sk[3 * (size_t)idx + 2] = pek;
  } else if (idx < a.ncube) {
    sk[3 * (size_t)idx + 2] = 0.0;  // outside the sphere: no term
  }
  const double w = wave_sum(e);
  if ((threadIdx.x & 63) == 0) she[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (she[0] + she[1]) + (she[2] + she[3]);
}

// forces -dE/dr_i (CpuEwaldRecip.cpp:125-139): one thread per particle, the
// k-vectors in the reference's order, staged through LDS
__global__ __launch_bounds__(kErBlock) void k_er_force(ErArgs a, const P4<double>* __restrict__ xyzq,
                                                       int n, const double* __restrict__ sk,
                                                       double* __restrict__ fx,
                                                       double* __restrict__ fy,
                                                       double* __restrict__ fz) {
  __shared__ double shk[kErBlock][6];
  const int i = blockIdx.x * kErBlock + threadIdx.x;
  P4<double> p{0, 0, 0, 0};
  if (i < n) p = xyzq[i];
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int base = 0; base < a.ncube; base += kErBlock) {
    __syncthreads();
    const int k = base + threadIdx.x;
    if (k < a.ncube) {
      double kv[3];
      const bool ok = er_kvec(a, k, kv);
      shk[threadIdx.x][0] = kv[0];
      shk[threadIdx.x][1] = kv[1];
      shk[threadIdx.x][2] = kv[2];
      shk[threadIdx.x][3] = ok ? sk[3 * (size_t)k + 0] : 0.0;
      shk[threadIdx.x][4] = ok ? sk[3 * (size_t)k + 1] : 0.0;
      shk[threadIdx.x][5] = ok ? sk[3 * (size_t)k + 2] : 0.0;
    }
    __syncthreads();
    const int m = min(kErBlock, a.ncube - base);
    if (i < n)
      for (int j = 0; j < m; j++) {
        const double pek = shk[j][5];
        if (pek == 0.0) continue;
        const double kx = shk[j][0], ky = shk[j][1], kz = shk[j][2];
        const double re = shk[j][3], im = shk[j][4];
        const double kr = kx * p.x + ky * p.y + kz * p.z;
        double sn, cs;
        sincos(kr, &sn, &cs);
        const double fac = -pek * p.w * 2.0;
        f0 += fac * (re * (-kx * sn) + im * kx * cs);
        f1 += fac * (re * (-ky * sn) + im * ky * cs);
        f2 += fac * (re * (-kz * sn) + im * kz * cs);
      }
  }
  if (i < n) {
    fx[i] = f0;
    fy[i] = f1;
    fz[i] = f2;
  }
}

__global__ void k_er_finalize(const double* __restrict__ partials, int nb, double* energy) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partials[b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) energy[0] = sh[0];
}

int ewald_recip_blocks(const ErArgs& a) { return (a.ncube + kErBlock - 1) / kErBlock; }

hipError_t launch_ewald_recip(const ErArgs& a, const void* xyzq, int n, double* sk,
                              double* partials, double* fx, double* fy, double* fz,
                              double* energy, hipStream_t s) {
  const int nb = ewald_recip_blocks(a);
  k_er_sfac<<<nb, kErBlock, 0, s>>>(a, (const P4<double>*)xyzq, n, sk, partials);
  k_er_finalize<<<1, 256, 0, s>>>(partials, nb, energy);
  if (n > 0 && fx)
    k_er_force<<<(n + kErBlock - 1) / kErBlock, kErBlock, 0, s>>>(a, (const P4<double>*)xyzq, n,
                                                                  sk, fx, fy, fz);
  return hipGetLastError();
}

}  // namespace gse
