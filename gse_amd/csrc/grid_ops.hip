// k-space pass, reductions and grid operators.
//
//  k_greens_table  G(k) in FP64 exactly as CpuGreensFuncGSE.cpp:33-56
//                  (kind 0; sigma = 0 gives 4 pi / k^2) and
//                  CpuGreensFuncG2.cpp:41-60 (kind 1), stored in the grid type
//  k_kpass         one pass over the r2c half spectrum: multiply by G and
//                  accumulate the energy sum_k mult G |rho_k|^2 (Parseval form
//                  of CpuGSE.cpp:86-104) and, optionally, the six k-space
//                  virial moments, all in FP64, per block
//  k_dot / k_sum_columns / k_reduce_partials   deterministic FP64 reductions
//                  (fixed block count, fixed combination order)
//  k_conv          one separable periodic convolution pass (gather form,
//                  CpuGaussCharge.cpp:448-514 / CpuConvolution.cpp:134-148);
//                  FP32 sums are compensated (TwoProd/TwoSum)
//  k_efield        CpuGaussCharge.cpp:62-99
//  k_dipole        CpuGaussCharge.cpp:189-226
#include <math.h>

#include "gse_device.h"
#include "gse_internal.h"

namespace gse {

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};
template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

// Block-wide sum of NC doubles; thread 0 writes partials[blockIdx.x*NC + c].
template <int NC>
__device__ __forceinline__ void block_sum_store(double (&v)[NC], double* partials) {
  __shared__ double sh[NC][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    double s = wave_sum(v[c]);
    if (lane == 0) sh[c][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh[threadIdx.x][w];
    partials[blockIdx.x * NC + threadIdx.x] = s;
  }
}

struct KArgs {
  int X, Y, Z, Xh;
  int halfX, halfY, halfZ;
  double twopi_bx, twopi_by, twopi_bz;
  double fourpi_vol, inv_vol, sigmafac;
};

static KArgs make_kargs(int X, int Y, int Z, double bx, double by, double bz, double sigma) {
  KArgs a;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.Xh = X / 2 + 1;
  a.halfX = X / 2 + (X % 2);
  a.halfY = Y / 2 + (Y % 2);
  a.halfZ = Z / 2 + (Z % 2);
  a.twopi_bx = 2.0 * M_PI / bx;
  a.twopi_by = 2.0 * M_PI / by;
  a.twopi_bz = 2.0 * M_PI / bz;
  a.fourpi_vol = 4.0 * M_PI / (double)(X * Y * Z);
  a.inv_vol = 1.0 / (double)(X * Y * Z);
  a.sigmafac = sigma * sigma / 2.0;
  return a;
}

__device__ __forceinline__ void kvec(const KArgs& a, size_t idx, int& ix, double& kx, double& ky,
                                     double& kz) {
  ix = (int)(idx % a.Xh);
  const size_t t = idx / a.Xh;
  const int iy = (int)(t % a.Y);
  const int iz = (int)(t / a.Y);
  const int mx = (ix >= a.halfX) ? ix - a.X : ix;
  const int my = (iy >= a.halfY) ? iy - a.Y : iy;
  const int mz = (iz >= a.halfZ) ? iz - a.Z : iz;
  kx = mul_rn((double)mx, a.twopi_bx);
  ky = mul_rn((double)my, a.twopi_by);
  kz = mul_rn((double)mz, a.twopi_bz);
}

template <typename T>
__global__ void k_greens_table(KArgs a, int kind, T* __restrict__ table) {
  const size_t n = (size_t)a.Xh * a.Y * a.Z;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    int ix;
    double kx, ky, kz;
    kvec(a, idx, ix, kx, ky, kz);
    const double kd2 = add_rn(add_rn(mul_rn(kx, kx), mul_rn(ky, ky)), mul_rn(kz, kz));
    double fac;
    if (kind == 0) {
      fac = idx == 0 ? 0.0 : a.fourpi_vol * exp(-a.sigmafac * kd2) / kd2;
    } else {
      fac = a.inv_vol * exp(-a.sigmafac * kd2);
    }
    table[idx] = (T)fac;
  }
}

template <typename T, bool VIR>
__global__ void __launch_bounds__(256)
    k_kpass(KArgs a, int kind, C2<T>* __restrict__ spec, const T* __restrict__ table,
            double* __restrict__ partials) {
  const size_t n = (size_t)a.Xh * a.Y * a.Z;
  double e = 0.0, vxx = 0.0, vxy = 0.0, vxz = 0.0, vyy = 0.0, vyz = 0.0, vzz = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    C2<T> c = spec[idx];
    const T g = table[idx];
    const double re = (double)c.x, im = (double)c.y;
    c.x = c.x * g;
    c.y = c.y * g;
    spec[idx] = c;
    const int ix = (int)(idx % a.Xh);
    // half-spectrum weight: columns 0 and X/2 (even X) stand for themselves
    const double mult = (ix == 0 || 2 * ix == a.X) ? 1.0 : 2.0;
    const double ek = mult * (double)g * (re * re + im * im);
    e += ek;
    if constexpr (VIR) {
      int ixx;
      double kx, ky, kz;
      kvec(a, idx, ixx, kx, ky, kz);
      const double k2 = kx * kx + ky * ky + kz * kz;
      if (k2 > 0.0) {
        // dG/d(eps_ab) = 2 k_a k_b G (s + 1/k^2) (no 1/k^2 term for G2)
        const double cf = ek * 2.0 * (a.sigmafac + (kind == 0 ? 1.0 / k2 : 0.0));
        vxx += cf * kx * kx;
        vxy += cf * kx * ky;
        vxz += cf * kx * kz;
        vyy += cf * ky * ky;
        vyz += cf * ky * kz;
        vzz += cf * kz * kz;
      }
    }
  }
  if constexpr (VIR) {
    double acc[7] = {e, vxx, vxy, vxz, vyy, vyz, vzz};
    block_sum_store<7>(acc, partials);
  } else {
    double acc[1] = {e};
    block_sum_store<1>(acc, partials);
  }
}

__global__ void __launch_bounds__(256)
    k_kpass_finalize(const double* __restrict__ partials, int nblocks, int ncomp, double pref,
                     double* __restrict__ energy, double* __restrict__ virial) {
  kpass_finalize_block(partials, nblocks, ncomp, pref, energy, virial);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_dot(const T* __restrict__ a, const T* __restrict__ b, size_t n, double* partials) {
  double acc[1] = {0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    acc[0] += (double)(a[i] * b[i]);
  block_sum_store<1>(acc, partials);
}

__global__ void __launch_bounds__(256)
    k_sum_columns(const double* __restrict__ v, int n, int ncomp, double* partials) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < ncomp && c < 6; c++) acc[c] += v[(size_t)c * n + i];
  block_sum_store<6>(acc, partials);
}

template <int NC>
__global__ void __launch_bounds__(256)
    k_reduce_partials(const double* __restrict__ partials, int nblocks, int ncomp, int stride,
                      double scale, double* __restrict__ out) {
  double t[NC];
  reduce_partials_block<NC>(partials, nblocks, stride, t);
  if (threadIdx.x < ncomp && threadIdx.x < NC) out[threadIdx.x] = t[threadIdx.x] * scale;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_charge_moments(const V4<T>* __restrict__ xyzq, int n, double* partials) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double q = (double)xyzq[i].w;
    acc[0] += q;
    acc[1] += q * q;
  }
  block_sum_store<2>(acc, partials);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_dipole(int X, int Y, int Z, T hx, T hy, T hz, const T* __restrict__ rho, double* partials) {
  double acc[3] = {0, 0, 0};
  const size_t n = (size_t)X * Y * Z;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(p % X);
    const size_t t = p / X;
    const int iy = (int)(t % Y), iz = (int)(t / Y);
    const T val = rho[p] * hx * hy * hz;
    acc[0] += (double)(val * (hx * (T)ix));
    acc[1] += (double)(val * (hy * (T)iy));
    acc[2] += (double)(val * (hz * (T)iz));
  }
  block_sum_store<3>(acc, partials);
}

template <typename T>
__global__ void k_efield(int X, int Y, int Z, T ihx, T ihy, T ihz, const T* __restrict__ phi,
                         T* __restrict__ ex, T* __restrict__ ey, T* __restrict__ ez) {
  const size_t n = (size_t)X * Y * Z;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(p % X);
    const size_t t = p / X;
    const int iy = (int)(t % Y), iz = (int)(t / Y);
    const T c = phi[p];
    const T px = phi[(ix + 1 == X ? 0 : ix + 1) + (size_t)X * (iy + (size_t)Y * iz)];
    const T py = phi[ix + (size_t)X * ((iy + 1 == Y ? 0 : iy + 1) + (size_t)Y * iz)];
    const T pz = phi[ix + (size_t)X * (iy + (size_t)Y * (iz + 1 == Z ? 0 : iz + 1))];
    ex[p] = -(px - c) * ihx;
    ey[p] = -(py - c) * ihy;
    ez[p] = -(pz - c) * ihz;
  }
}

// One separable pass in gather form: out[i] = sum_{ii=-R..R} k[ii+R] in[i+ii]
// along the pass axis, periodic, summed in ii order.  Threads run x-fastest,
// so loads are coalesced on every axis.
template <typename T, int PASS>
__global__ void __launch_bounds__(256)
    k_conv(int X, int Y, int Z, int R, const T* __restrict__ taps, const T* __restrict__ in,
           T* __restrict__ out) {
  const size_t n = (size_t)X * Y * Z;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(p % X);
    const size_t t = p / X;
    const int iy = (int)(t % Y), iz = (int)(t / Y);
    const int M = PASS == 0 ? X : PASS == 1 ? Y : Z;
    const int c = PASS == 0 ? ix : PASS == 1 ? iy : iz;
    const size_t stride = PASS == 0 ? 1 : PASS == 1 ? (size_t)X : (size_t)X * Y;
    const size_t base = p - (size_t)c * stride;
    if (sizeof(T) == 8) {
      T s = 0;
      for (int ii = -R; ii <= R; ii++) {
        const int q = wrapi(c + ii, M);
        s += taps[ii + R] * in[base + (size_t)q * stride];
      }
      out[p] = s;
    } else {
      KSum<T> s;
      for (int ii = -R; ii <= R; ii++) {
        const int q = wrapi(c + ii, M);
        s.add_prod(taps[ii + R], in[base + (size_t)q * stride]);
      }
      out[p] = s.value();
    }
  }
}

// ---------------------------------------------------------------------------
static unsigned grid_for(size_t count, unsigned threads, unsigned cap) {
  size_t g = (count + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

hipError_t launch_greens_table(bool dbl, int kind, int X, int Y, int Z, double bx, double by,
                               double bz, double sigma, void* table, hipStream_t s) {
  const KArgs a = make_kargs(X, Y, Z, bx, by, bz, sigma);
  const size_t n = (size_t)a.Xh * Y * Z;
  if (dbl)
    k_greens_table<double><<<grid_for(n, 256, 8192), 256, 0, s>>>(a, kind, (double*)table);
  else
    k_greens_table<float><<<grid_for(n, 256, 8192), 256, 0, s>>>(a, kind, (float*)table);
  return hipGetLastError();
}

hipError_t launch_kpass(bool dbl, int kind, int X, int Y, int Z, double bx, double by, double bz,
                        double sigma, void* spec, const void* table, bool vir, double* partials,
                        hipStream_t s) {
  const KArgs a = make_kargs(X, Y, Z, bx, by, bz, sigma);
  if (dbl) {
    if (vir)
      k_kpass<double, true><<<kKpassBlocks, 256, 0, s>>>(a, kind, (C2<double>*)spec,
                                                         (const double*)table, partials);
    else
      k_kpass<double, false><<<kKpassBlocks, 256, 0, s>>>(a, kind, (C2<double>*)spec,
                                                          (const double*)table, partials);
  } else {
    if (vir)
      k_kpass<float, true><<<kKpassBlocks, 256, 0, s>>>(a, kind, (C2<float>*)spec,
                                                        (const float*)table, partials);
    else
      k_kpass<float, false><<<kKpassBlocks, 256, 0, s>>>(a, kind, (C2<float>*)spec,
                                                         (const float*)table, partials);
  }
  return hipGetLastError();
}

hipError_t launch_kpass_finalize(const double* partials, int ncomp, double pref, double* energy,
                                 double* virial, hipStream_t s) {
  k_kpass_finalize<<<1, 256, 0, s>>>(partials, kKpassBlocks, ncomp, pref, energy, virial);
  return hipGetLastError();
}

hipError_t launch_dot(bool dbl, const void* a, const void* b, size_t n, double* partials,
                      hipStream_t s) {
  if (dbl)
    k_dot<double><<<kRedBlocks, 256, 0, s>>>((const double*)a, (const double*)b, n, partials);
  else
    k_dot<float><<<kRedBlocks, 256, 0, s>>>((const float*)a, (const float*)b, n, partials);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partials, int nblocks, int ncomp, int stride,
                                  double scale, double* out, hipStream_t s) {
  if (ncomp <= 1)
    k_reduce_partials<1><<<1, 256, 0, s>>>(partials, nblocks, ncomp, stride, scale, out);
  else if (ncomp <= 3)
    k_reduce_partials<3><<<1, 256, 0, s>>>(partials, nblocks, ncomp, stride, scale, out);
  else
    k_reduce_partials<6><<<1, 256, 0, s>>>(partials, nblocks, ncomp, stride, scale, out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
    k_add_sym6(const double* __restrict__ partials, int nblocks, double* vir9) {
  double t[6];
  reduce_partials_block<6>(partials, nblocks, 6, t);
  if (threadIdx.x < 9) {
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    vir9[threadIdx.x] += t[map[threadIdx.x]];
  }
}

hipError_t launch_add_sym6(const double* partials, int nblocks, double* vir9, hipStream_t s) {
  k_add_sym6<<<1, 256, 0, s>>>(partials, nblocks, vir9);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
    k_corrections(const double* __restrict__ partials, int nblocks, double sigma, double vol,
                  double ccelec, double* out) {
  double t[2];
  reduce_partials_block<2>(partials, nblocks, 2, t);
  if (threadIdx.x == 0) {
    out[0] = -ccelec * t[1] / sqrt(2.0 * M_PI * sigma * sigma);
    out[1] = -ccelec * M_PI * t[0] * t[0] * sigma * sigma / vol;
  }
}

hipError_t launch_corrections_finalize(const double* partials, int nblocks, double sigma,
                                       double bx, double by, double bz, double ccelec,
                                       double* out, hipStream_t s) {
  k_corrections<<<1, 256, 0, s>>>(partials, nblocks, sigma, bx * by * bz, ccelec, out);
  return hipGetLastError();
}

hipError_t launch_sum_columns(const double* v, int n, int ncomp, double* partials,
                              hipStream_t s) {
  k_sum_columns<<<kRedBlocks, 256, 0, s>>>(v, n, ncomp, partials);
  return hipGetLastError();
}

hipError_t launch_charge_moments(bool dbl, const void* xyzq, int n, double* partials,
                                 hipStream_t s) {
  if (dbl)
    k_charge_moments<double><<<kRedBlocks, 256, 0, s>>>((const V4<double>*)xyzq, n, partials);
  else
    k_charge_moments<float><<<kRedBlocks, 256, 0, s>>>((const V4<float>*)xyzq, n, partials);
  return hipGetLastError();
}

hipError_t launch_dipole(bool dbl, int X, int Y, int Z, double bx, double by, double bz,
                         const void* rho, double* partials, hipStream_t s) {
  const double hx = bx / X, hy = by / Y, hz = bz / Z;
  if (dbl)
    k_dipole<double><<<kRedBlocks, 256, 0, s>>>(X, Y, Z, hx, hy, hz, (const double*)rho, partials);
  else
    k_dipole<float><<<kRedBlocks, 256, 0, s>>>(X, Y, Z, (float)hx, (float)hy, (float)hz,
                                               (const float*)rho, partials);
  return hipGetLastError();
}

hipError_t launch_efield(bool dbl, int X, int Y, int Z, double bx, double by, double bz,
                         const void* phi, void* ex, void* ey, void* ez, hipStream_t s) {
  const size_t n = (size_t)X * Y * Z;
  const double ihx = 1.0 / (bx / X), ihy = 1.0 / (by / Y), ihz = 1.0 / (bz / Z);
  if (dbl)
    k_efield<double><<<grid_for(n, 256, 8192), 256, 0, s>>>(X, Y, Z, ihx, ihy, ihz,
                                                            (const double*)phi, (double*)ex,
                                                            (double*)ey, (double*)ez);
  else
    k_efield<float><<<grid_for(n, 256, 8192), 256, 0, s>>>(X, Y, Z, (float)ihx, (float)ihy,
                                                           (float)ihz, (const float*)phi,
                                                           (float*)ex, (float*)ey, (float*)ez);
  return hipGetLastError();
}

hipError_t launch_conv_pass(bool dbl, int pass, int X, int Y, int Z, int R, const void* taps,
                            const void* in, void* out, hipStream_t s) {
  const size_t n = (size_t)X * Y * Z;
  const unsigned g = grid_for(n, 256, 16384);
#define GSE_CONV(T, P) \
  k_conv<T, P><<<g, 256, 0, s>>>(X, Y, Z, R, (const T*)taps, (const T*)in, (T*)out)
  if (dbl) {
    if (pass == 0) GSE_CONV(double, 0);
    else if (pass == 1) GSE_CONV(double, 1);
    else GSE_CONV(double, 2);
  } else {
    if (pass == 0) GSE_CONV(float, 0);
    else if (pass == 1) GSE_CONV(float, 1);
    else GSE_CONV(float, 2);
  }
#undef GSE_CONV
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Batched independent systems: the k-pass over every system's half spectrum
// (one chunk of one system per workgroup, FP64 partial per chunk) and the
// per-system energy E_s = 1/2 ccelec h^3 sum_k w G |rho_k|^2.
template <typename T>
__global__ void __launch_bounds__(256)
    k_batch_kpass(BatchDev d, const BatchChunk* __restrict__ chunks, C2<T>* __restrict__ spec,
                  const T* __restrict__ table, double* __restrict__ partials) {
  const BatchChunk c = chunks[blockIdx.x];
  const BatchSys& S = d.sys[c.sys];
  const int X = S.g.X, Xh = X / 2 + 1;
  C2<T>* sp = spec + S.kofs;
  const T* tb = table + S.kofs;
  double e = 0.0;
  for (int i = c.start + threadIdx.x; i < c.start + c.count; i += blockDim.x) {
    C2<T> v = sp[i];
    const T g = tb[i];
    const double re = (double)v.x, im = (double)v.y;
    v.x = v.x * g;
    v.y = v.y * g;
    sp[i] = v;
    const int ix = i % Xh;
    const double mult = (ix == 0 || 2 * ix == X) ? 1.0 : 2.0;
    e += mult * (double)g * (re * re + im * im);
  }
  double acc[1] = {e};
  block_sum_store<1>(acc, partials);
}

__global__ void __launch_bounds__(64)
    k_batch_energy(BatchDev d, const double* __restrict__ partials, double ccelec,
                   double* __restrict__ energies) {
  const BatchSys& S = d.sys[blockIdx.x];
  double v = 0.0;
  for (int b = threadIdx.x; b < S.nkblk; b += 64) v += partials[S.kblk0 + b];
  v = wave_sum(v);
  if (threadIdx.x == 0) energies[blockIdx.x] = 0.5 * ccelec * S.vcell * v;
}

hipError_t launch_batch_kpass(bool dbl, const BatchDev& d, const BatchChunk* chunks, int nchunks,
                              void* spec, const void* table, double* partials, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  if (dbl)
    k_batch_kpass<double><<<nchunks, 256, 0, s>>>(d, chunks, (C2<double>*)spec,
                                                  (const double*)table, partials);
  else
    k_batch_kpass<float><<<nchunks, 256, 0, s>>>(d, chunks, (C2<float>*)spec, (const float*)table,
                                                 partials);
  return hipGetLastError();
}

hipError_t launch_batch_energy(const BatchDev& d, const double* partials, double ccelec,
                               double* energies, hipStream_t s) {
  if (d.nsys <= 0) return hipSuccess;
  k_batch_energy<<<d.nsys, 64, 0, s>>>(d, partials, ccelec, energies);
  return hipGetLastError();
}

}  // namespace gse
