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
//  k_conv_x4 / k_conv_x / k_conv_pencil / k_conv_yz  one separable periodic
//                  convolution pass (gather form, CpuGaussCharge.cpp:448-514 /
//                  CpuConvolution.cpp:134-148); FP32 sums are Kahan-compensated
//  k_efield        CpuGaussCharge.cpp:62-99
//  k_dipole        CpuGaussCharge.cpp:189-226
#include <math.h>

#include <stdint.h>

#include <algorithm>

#include "gse_device.h"
#include "gse_internal.h"
#include "zdft_device.h"

namespace gse {

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
  int zs;  // elements between consecutive z planes of the spectrum / G table
  int halfX, halfY, halfZ;
  double twopi_bx, twopi_by, twopi_bz;
  double fourpi_vol, inv_vol, sigmafac;
  double ih2[3];  // (M_d / L_d)^2, the HERMITIAN stencil's 1/h^2
};

static KArgs make_kargs(int X, int Y, int Z, double bx, double by, double bz, double sigma) {
  KArgs a;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.Xh = X / 2 + 1;
  a.zs = a.Xh * Y;
  a.halfX = X / 2 + (X % 2);
  a.halfY = Y / 2 + (Y % 2);
  a.halfZ = Z / 2 + (Z % 2);
  a.twopi_bx = 2.0 * M_PI / bx;
  a.twopi_by = 2.0 * M_PI / by;
  a.twopi_bz = 2.0 * M_PI / bz;
  a.fourpi_vol = 4.0 * M_PI / (double)(X * Y * Z);
  a.inv_vol = 1.0 / (double)(X * Y * Z);
  a.sigmafac = sigma * sigma / 2.0;
  const double ih[3] = {(double)X / bx, (double)Y / by, (double)Z / bz};
  for (int d = 0; d < 3; d++) a.ih2[d] = ih[d] * ih[d];
  return a;
}

// The HERMITIAN (Mehrstellen) discretisation of CpuMultiGridSolver.cpp:
// 237-251, A phi = B rho' with rho' = -4 pi rhoS, solved exactly per wave
// vector: -4 pi B^/(A^ XYZ), zero at k = 0.  With s_d = 2 sin^2(pi i_d/M_d)
// the symbols lose no digits near k = 0:
//   A^ = -2 sum_d b_d s_d - 4 sum_{d<e} c_de (s_d + s_e - s_d s_e)
//   B^ = -1 + (s_x + s_y + s_z)/6
__device__ __forceinline__ double mehrstellen_g(const KArgs& a, size_t idx) {
  if (idx == 0) return 0.0;
  const int ix = (int)(idx % a.Xh);
  const size_t t = idx / a.Xh;
  const int iy = (int)(t % a.Y), iz = (int)(t / a.Y);
  double sx = sin(M_PI * ix / a.X), sy = sin(M_PI * iy / a.Y), sz = sin(M_PI * iz / a.Z);
  sx = 2.0 * sx * sx;
  sy = 2.0 * sy * sy;
  sz = 2.0 * sz * sz;
  const double S = a.ih2[0] + a.ih2[1] + a.ih2[2];
  const double bx = -5.0 / 6.0 * a.ih2[0] + S / 6.0, by = -5.0 / 6.0 * a.ih2[1] + S / 6.0,
               bz = -5.0 / 6.0 * a.ih2[2] + S / 6.0;
  const double cxy = -(a.ih2[0] + a.ih2[1]) / 12.0, cxz = -(a.ih2[0] + a.ih2[2]) / 12.0,
               cyz = -(a.ih2[1] + a.ih2[2]) / 12.0;
  const double A = -2.0 * (bx * sx + by * sy + bz * sz) -
                   4.0 * (cxy * (sx + sy - sx * sy) + cxz * (sx + sz - sx * sz) +
                          cyz * (sy + sz - sy * sz));
  const double B = -1.0 + (sx + sy + sz) / 6.0;
  return -4.0 * M_PI * B / (A * (double)a.X * a.Y * a.Z);
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

// table[z zs + y Xh + x] (zs >= Xh Y: a padded plane stride leaves the
// entries past Xh Y of each plane unwritten)
template <typename T>
__global__ void k_greens_table(KArgs a, int kind, T* __restrict__ table) {
  const size_t plane = (size_t)a.Xh * a.Y;
  const size_t n = plane * a.Z;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    int ix;
    double kx, ky, kz;
    kvec(a, idx, ix, kx, ky, kz);
    const double kd2 = add_rn(add_rn(mul_rn(kx, kx), mul_rn(ky, ky)), mul_rn(kz, kz));
    double fac;
    if (kind == 0) {
      fac = idx == 0 ? 0.0 : a.fourpi_vol * exp(-a.sigmafac * kd2) / kd2;
    } else if (kind == 1) {
      fac = a.inv_vol * exp(-a.sigmafac * kd2);
    } else {
      fac = mehrstellen_g(a, idx);
    }
    table[(idx / plane) * a.zs + idx % plane] = (T)fac;
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

// ---------------------------------------------------------------------------
// Fused z pass: between a 2D r2c and a 2D c2r over the x-y planes, each
// one-wave workgroup takes C consecutive columns (ky, kx) of the half
// spectrum [z][y][x <= X/2] and, in LDS, runs the z transform (mixed-radix
// Stockham, radices 2..8, twiddles exp(-2 pi i m/Z) from a table), the G(k)
// multiply with the FP64 energy / virial partial sums of k_kpass, and the
// inverse z transform, then stores the columns in place.  Stockham step of
// radix p after a sub-length T (k = i mod T, j = (i - k) p + k):
//   y_s = sum_q (x[i + q Z/p] w^(q k Z/(p T))) w^(q s Z/p),  x'[j + s T] = y_s
// The p roots w^(r Z/p) of the inner DFT are read once per stage into
// registers; q k Z/(p T) < Z, so the stage twiddle needs no reduction.
// ---------------------------------------------------------------------------

template <typename T, int P, int C>
__device__ __forceinline__ void zfft_butterfly(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst,
                                               const C2<T>* __restrict__ W, int Z, int Tsub,
                                               bool inv) {
  const int m = Z / P;
  const int wstep = Z / (P * Tsub);
  C2<T> root[P];
#pragma unroll
  for (int r = 0; r < P; r++) {
    root[r] = W[r * m];
    if (inv) root[r].y = -root[r].y;
  }
  for (int w = threadIdx.x; w < m * C; w += blockDim.x) {
    const int i = w / C, cc = w % C;  // C is a power of two
    const int k = i % Tsub;
    C2<T> x[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
      C2<T> v = src[(i + q * m) * C + cc];
      if (q > 0 && k > 0) {
        C2<T> t = W[q * k * wstep];
        if (inv) t.y = -t.y;
        v = C2<T>{v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x};
      }
      x[q] = v;
    }
    const int j = (i - k) * P + k;
    if (zdft_small<T, P>(x, inv)) {
#pragma unroll
      for (int s = 0; s < P; s++) dst[(j + s * Tsub) * C + cc] = x[s];
    } else {
#pragma unroll
      for (int s = 0; s < P; s++) {
        T yr = x[0].x, yi = x[0].y;
#pragma unroll
        for (int q = 1; q < P; q++) {
          const C2<T> t = root[(q * s) % P];
          yr += x[q].x * t.x - x[q].y * t.y;
          yi += x[q].x * t.y + x[q].y * t.x;
        }
        dst[(j + s * Tsub) * C + cc] = C2<T>{yr, yi};
      }
    }
  }
}

// the whole z transform; returns the buffer holding the result
template <typename T, int C>
__device__ C2<T>* zfft(C2<T>* a, C2<T>* b, const C2<T>* W, const ZPlan& pl, bool inv) {
  int Tsub = 1;
  for (int st = 0; st < pl.nst; st++) {
    const int p = pl.p[st];
    switch (p) {
      case 2: zfft_butterfly<T, 2, C>(a, b, W, pl.Z, Tsub, inv); break;
      case 3: zfft_butterfly<T, 3, C>(a, b, W, pl.Z, Tsub, inv); break;
      case 4: zfft_butterfly<T, 4, C>(a, b, W, pl.Z, Tsub, inv); break;
      case 5: zfft_butterfly<T, 5, C>(a, b, W, pl.Z, Tsub, inv); break;
      case 6: zfft_butterfly<T, 6, C>(a, b, W, pl.Z, Tsub, inv); break;
      case 7: zfft_butterfly<T, 7, C>(a, b, W, pl.Z, Tsub, inv); break;
      default: zfft_butterfly<T, 8, C>(a, b, W, pl.Z, Tsub, inv); break;
    }
    __syncthreads();
    C2<T>* t = a;
    a = b;
    b = t;
    Tsub *= p;
  }
  return a;
}

constexpr int kZpassThreads = 256;
constexpr int kZpassPerThread = 8;  // Z * C <= 2048 complex values per workgroup

template <typename T, int C, bool VIR>
__global__ void __launch_bounds__(kZpassThreads)
    k_zpass(KArgs a, ZPlan pl, int kind, C2<T>* __restrict__ spec, const T* __restrict__ table,
            const C2<T>* __restrict__ tw, double* __restrict__ partials) {
  extern __shared__ __align__(16) char zsm[];
  const int Z = pl.Z, total = Z * C;
  C2<T>* A = (C2<T>*)zsm;
  C2<T>* B = A + (size_t)Z * C;
  C2<T>* W = B + (size_t)Z * C;
  const int ncol = a.Xh * a.Y;
  const size_t zs = (size_t)a.zs;  // memory stride of z (kvec takes ncol)
  const int col0 = blockIdx.x * C;
  // every load of the workgroup's columns and of their G(k) values is issued
  // before the first LDS store, so a thread has all of them in flight; the
  // G values stay in registers for the multiply (same item per thread)
  C2<T> v[kZpassPerThread];
  T gk[kZpassPerThread];
#pragma unroll
  for (int u = 0; u < kZpassPerThread; u++) {
    const int idx = threadIdx.x + u * kZpassThreads;
    const int z = idx / C, c = col0 + idx % C;
    const bool ok = idx < total && c < ncol;
    const size_t kidx = ok ? z * zs + c : 0;
    v[u] = ok ? spec[kidx] : C2<T>{(T)0, (T)0};
    gk[u] = ok ? table[kidx] : (T)0;
  }
  for (int i = threadIdx.x; i < Z; i += kZpassThreads) W[i] = tw[i];
#pragma unroll
  for (int u = 0; u < kZpassPerThread; u++) {
    const int idx = threadIdx.x + u * kZpassThreads;
    if (idx < total) A[idx] = v[u];
  }
  __syncthreads();
  C2<T>* F = zfft<T, C>(A, B, W, pl, false);
  C2<T>* G = F == A ? B : A;
  double e = 0.0, vxx = 0.0, vxy = 0.0, vxz = 0.0, vyy = 0.0, vyz = 0.0, vzz = 0.0;
#pragma unroll
  for (int u = 0; u < kZpassPerThread; u++) {
    const int idx = threadIdx.x + u * kZpassThreads;
    const int z = idx / C, c = col0 + idx % C;
    if (idx >= total || c >= ncol) continue;
    const C2<T> f = F[idx];
    const T g = gk[u];
    const double re = (double)f.x, im = (double)f.y;
    F[idx] = C2<T>{f.x * g, f.y * g};
    const int ix = c % a.Xh;
    const double mult = (ix == 0 || 2 * ix == a.X) ? 1.0 : 2.0;
    const double ek = mult * (double)g * (re * re + im * im);
    e += ek;
    if constexpr (VIR) {
      int ixx;
      double kx, ky, kz;
      kvec(a, (size_t)z * ncol + c, ixx, kx, ky, kz);
      const double k2 = kx * kx + ky * ky + kz * kz;
      if (k2 > 0.0) {
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
  __syncthreads();
  C2<T>* R = zfft<T, C>(F, G, W, pl, true);
#pragma unroll
  for (int u = 0; u < kZpassPerThread; u++) {
    const int idx = threadIdx.x + u * kZpassThreads;
    const int z = idx / C, c = col0 + idx % C;
    if (idx < total && c < ncol) spec[z * zs + c] = R[idx];
  }
  if constexpr (VIR) {
    double acc[7] = {e, vxx, vxy, vxz, vyy, vyz, vzz};
    block_sum_store<7>(acc, partials);
  } else {
    double acc[1] = {e};
    block_sum_store<1>(acc, partials);
  }
}

// ---------------------------------------------------------------------------
// Four-step z pass for the column lengths it is built for (Z = Z1 Z2 known at
// compile time).  With z = Z2 n1 + n2 and k = k1 + Z1 k2,
//   X[k1 + Z1 k2] = sum_n2 w_Z2^(n2 k2) w_Z^(n2 k1) sum_n1 x[Z2 n1 + n2] w_Z1^(n1 k1):
// thread n2 of a column holds the Z1 values of its stride-Z2 sub-column and
// transforms them in registers, multiplies by w_Z^(n2 k1) and writes them to
// LDS; after one barrier thread k1 reads the Z2 values of row k1 and
// transforms them in registers.  The G(k) multiply and the FP64 energy /
// virial sums follow in registers, then the inverse runs the two steps back
// with conjugate twiddles.  Two LDS exchanges per direction instead of one
// per Stockham stage; the in-register DFTs (reg_dft) take every twiddle as
// w_Z^j from the table.
// ---------------------------------------------------------------------------
template <int Z1, int Z2, int C>
constexpr int zpass4_threads() {
  return ((C * (Z1 > Z2 ? Z1 : Z2)) + 63) / 64 * 64;
}

template <typename T, int Z1, int Z2, int C, bool VIR>
__global__ void __launch_bounds__((zpass4_threads<Z1, Z2, C>()))
    k_zpass4(KArgs a, int kind, C2<T>* __restrict__ spec, const T* __restrict__ table,
             const C2<T>* __restrict__ tw, double* __restrict__ partials) {
  constexpr int Z = Z1 * Z2;
  constexpr int NT = zpass4_threads<Z1, Z2, C>();
  extern __shared__ __align__(16) char zsm[];
  C2<T>* buf = (C2<T>*)zsm;  // [Z][C]: entry (k1 Z2 + n2) of every column
  C2<T>* W = buf + (size_t)Z * C;
  const int ncol = a.Xh * a.Y;
  const size_t zs = (size_t)a.zs;  // memory stride of z (kvec takes ncol)
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  const bool okc = col < ncol;
  // every load first: the thread's sub-column (step one) and its G(k) row
  // (step two), all in flight at once
  C2<T> u[Z1];
  T g[Z2];
#pragma unroll
  for (int n1 = 0; n1 < Z1; n1++)
    u[n1] = (t < Z2 && okc) ? spec[(size_t)(Z2 * n1 + t) * zs + col] : C2<T>{(T)0, (T)0};
#pragma unroll
  for (int k2 = 0; k2 < Z2; k2++)
    g[k2] = (t < Z1 && okc) ? table[(size_t)(t + Z1 * k2) * zs + col] : (T)0;
  for (int i = threadIdx.x; i < Z; i += NT) W[i] = tw[i];
  __syncthreads();
  if (t < Z2) {  // step one, thread n2 = t
    reg_dft<T, Z1, Z>(u, W, false);
#pragma unroll
    for (int k1 = 0; k1 < Z1; k1++) {
      const C2<T> v = k1 > 0 ? zc_mul(u[k1], W[t * k1], false) : u[k1];
      buf[(k1 * Z2 + t) * C + c] = v;
    }
  }
  __syncthreads();
  double e = 0.0, vxx = 0.0, vxy = 0.0, vxz = 0.0, vyy = 0.0, vyz = 0.0, vzz = 0.0;
  C2<T> v[Z2];
  if (t < Z1) {  // step two, thread k1 = t: v[k2] = X[t + Z1 k2]
#pragma unroll
    for (int n2 = 0; n2 < Z2; n2++) v[n2] = buf[(t * Z2 + n2) * C + c];
    reg_dft<T, Z2, Z>(v, W, false);
    const int ix = col % a.Xh;
    const double mult = (ix == 0 || 2 * ix == a.X) ? 1.0 : 2.0;
#pragma unroll
    for (int k2 = 0; k2 < Z2; k2++) {
      const double re = (double)v[k2].x, im = (double)v[k2].y;
      v[k2] = C2<T>{v[k2].x * g[k2], v[k2].y * g[k2]};
      if (!okc) continue;
      const double ek = mult * (double)g[k2] * (re * re + im * im);
      e += ek;
      if constexpr (VIR) {
        int ixx;
        double kx, ky, kz;
        kvec(a, (size_t)(t + Z1 * k2) * ncol + col, ixx, kx, ky, kz);
        const double k2n = kx * kx + ky * ky + kz * kz;
        if (k2n > 0.0) {
          const double cf = ek * 2.0 * (a.sigmafac + (kind == 0 ? 1.0 / k2n : 0.0));
          vxx += cf * kx * kx;
          vxy += cf * kx * ky;
          vxz += cf * kx * kz;
          vyy += cf * ky * ky;
          vyz += cf * ky * kz;
          vzz += cf * kz * kz;
        }
      }
    }
    reg_dft<T, Z2, Z>(v, W, true);  // v[n2] = Y'[t][n2]
#pragma unroll
    for (int n2 = 1; n2 < Z2; n2++)
      if (t > 0) v[n2] = zc_mul(v[n2], W[n2 * t], true);
  }
  __syncthreads();  // every read of the forward rows is done
  if (t < Z1) {
#pragma unroll
    for (int n2 = 0; n2 < Z2; n2++) buf[(t * Z2 + n2) * C + c] = v[n2];
  }
  __syncthreads();
  if (t < Z2) {  // inverse step one, thread n2 = t
#pragma unroll
    for (int k1 = 0; k1 < Z1; k1++) u[k1] = buf[(k1 * Z2 + t) * C + c];
    reg_dft<T, Z1, Z>(u, W, true);
    if (okc) {
#pragma unroll
      for (int n1 = 0; n1 < Z1; n1++) spec[(size_t)(Z2 * n1 + t) * zs + col] = u[n1];
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

// ---------------------------------------------------------------------------
// Four-step y pass: the y transform of every (z, kx) column of the half
// spectrum, Y = Y1 Y2 in registers with one LDS exchange, as k_zpass4 does
// along z.  Forward: src packed [z][y][kx] (a batched 1D x r2c's output),
// dst planes dzs elements apart; inverse: the reverse (src planes szs
// apart, dst packed, for the 1D x c2r).  Column q = z Xh + kx; C consecutive
// columns per workgroup, so loads and stores run along kx.
// ---------------------------------------------------------------------------
template <typename T, int Y1, int Y2, int C, bool INV>
__global__ void __launch_bounds__((zpass4_threads<Y1, Y2, C>()))
    k_ypass4(int Xh, int Z, const C2<T>* __restrict__ src, int szs, C2<T>* __restrict__ dst,
             int dzs, const C2<T>* __restrict__ tw) {
  constexpr int Y = Y1 * Y2;
  constexpr int NT = zpass4_threads<Y1, Y2, C>();
  extern __shared__ __align__(16) char ysm[];
  C2<T>* buf = (C2<T>*)ysm;  // [Y][C]
  C2<T>* W = buf + (size_t)Y * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int q = blockIdx.x * C + c;
  const bool okc = q < Xh * Z;
  const int z = okc ? q / Xh : 0, kx = okc ? q - z * Xh : 0;
  const C2<T>* s0 = src + (size_t)z * szs + kx;
  C2<T>* d0 = dst + (size_t)z * dzs + kx;
  C2<T> u[Y1];
#pragma unroll
  for (int n1 = 0; n1 < Y1; n1++)
    u[n1] = (t < Y2 && okc) ? s0[(size_t)(Y2 * n1 + t) * Xh] : C2<T>{(T)0, (T)0};
  for (int i = threadIdx.x; i < Y; i += NT) W[i] = tw[i];
  __syncthreads();
  if (t < Y2) {  // step one, thread n2 = t
    reg_dft<T, Y1, Y>(u, W, INV);
#pragma unroll
    for (int k1 = 0; k1 < Y1; k1++) {
      const C2<T> v = k1 > 0 ? zc_mul(u[k1], W[t * k1], INV) : u[k1];
      buf[(k1 * Y2 + t) * C + c] = v;
    }
  }
  __syncthreads();
  if (t < Y1 && okc) {  // step two, thread k1 = t: v[k2] = X[t + Y1 k2]
    C2<T> v[Y2];
#pragma unroll
    for (int n2 = 0; n2 < Y2; n2++) v[n2] = buf[(t * Y2 + n2) * C + c];
    reg_dft<T, Y2, Y>(v, W, INV);
#pragma unroll
    for (int k2 = 0; k2 < Y2; k2++) d0[(size_t)(t + Y1 * k2) * Xh] = v[k2];
  }
}

// Y = 216 = 12 x 18 (32 columns per workgroup); Y = 80 = 8 x 10 (32
// columns in FP32, 16 in FP64), where 1D x rocFFT transforms and this pass
// measured below rocFFT's single 2D kernels: FP64 stage times r2c 13.0
// against 14.2, c2r 13.9 against 14.1 us; FP32 11.8 against 12.2 and 11.9
// against 12.3 us (8 or 16 columns: no gain in FP32)
bool ypass4_shape(int Y, bool) { return Y == 216 || Y == 80; }

template <int Y1, int Y2, int C>
static hipError_t ypass4_t(bool dbl, bool inv, int Xh, int Z, const void* src, int szs,
                           void* dst, int dzs, const void* tw, hipStream_t s) {
  constexpr int Y = Y1 * Y2;
  constexpr int NT = zpass4_threads<Y1, Y2, C>();
  const unsigned g = (unsigned)(((size_t)Xh * Z + C - 1) / C);
  const size_t lds = ((size_t)Y * C + Y) * (dbl ? 16 : 8);
#define GSE_YPASS(TT, INVV)                                                                \
  k_ypass4<TT, Y1, Y2, C, INVV><<<g, NT, lds, s>>>(Xh, Z, (const C2<TT>*)src, szs,         \
                                                   (C2<TT>*)dst, dzs, (const C2<TT>*)tw)
  if (dbl) {
    if (inv) GSE_YPASS(double, true);
    else GSE_YPASS(double, false);
  } else {
    if (inv) GSE_YPASS(float, true);
    else GSE_YPASS(float, false);
  }
#undef GSE_YPASS
  return hipGetLastError();
}

hipError_t launch_ypass4(bool dbl, bool inv, int X, int Y, int Z, const void* src, int szs,
                         void* dst, int dzs, const void* tw, hipStream_t s) {
  if (!ypass4_shape(Y, dbl)) return hipErrorInvalidValue;
  const int Xh = X / 2 + 1;
  if (szs < Xh * Y || dzs < Xh * Y) return hipErrorInvalidValue;
  if (Y == 80) {
    if (dbl) return ypass4_t<8, 10, 16>(dbl, inv, Xh, Z, src, szs, dst, dzs, tw, s);
    return ypass4_t<8, 10, 32>(dbl, inv, Xh, Z, src, szs, dst, dzs, tw, s);
  }
  return ypass4_t<12, 18, 32>(dbl, inv, Xh, Z, src, szs, dst, dzs, tw, s);
}

// the four-step forms built: Z -> (Z1, Z2), 16 or 32 columns per workgroup
struct Zpass4Shape {
  int Z, Z1, Z2;
};
constexpr Zpass4Shape kZpass4Shapes[] = {{80, 8, 10}, {216, 12, 18}};

static bool zpass4_shape(int Z, Zpass4Shape* out) {
  for (const Zpass4Shape& sh : kZpass4Shapes)
    if (sh.Z == Z) {
      *out = sh;
      return true;
    }
  return false;
}

template <typename T, int Z1, int Z2, int C>
static void zpass4_t(const KArgs& a, int kind, void* spec, const void* table, const void* tw,
                     bool vir, double* partials, unsigned g, hipStream_t s) {
  constexpr int NT = zpass4_threads<Z1, Z2, C>();
  const size_t lds = ((size_t)Z1 * Z2 * C + (size_t)Z1 * Z2) * sizeof(C2<T>);
  if (vir)
    k_zpass4<T, Z1, Z2, C, true><<<g, NT, lds, s>>>(a, kind, (C2<T>*)spec, (const T*)table,
                                                    (const C2<T>*)tw, partials);
  else
    k_zpass4<T, Z1, Z2, C, false><<<g, NT, lds, s>>>(a, kind, (C2<T>*)spec, (const T*)table,
                                                     (const C2<T>*)tw, partials);
}

bool radix_plan(int n, ZPlan* pl) {
  // greedy, largest radix first (8 x 3 x 3 x 3 at Z = 216 measured faster
  // than 6 x 6 x 6)
  pl->Z = n;
  pl->nst = 0;
  pl->C = 0;
  pl->four = 0;
  int m = n;
  while (m > 1) {
    int p = 0;
    for (int r : {8, 7, 6, 5, 4, 3, 2})
      if (m % r == 0) {
        p = r;
        break;
      }
    if (!p || pl->nst == kZMaxStages) return false;
    pl->p[pl->nst++] = p;
    m /= p;
  }
  return n >= 1;
}

bool zplan_make(int Z, bool dbl, ZPlan* pl) {
  pl->four = 0;
  Zpass4Shape sh;
  if (tuning().zpass4 && zpass4_shape(Z, &sh)) {
    pl->Z = Z;
    pl->nst = 0;
    // columns per workgroup: 32 measured faster for FP32 at Z = 216 (40.5 vs
    // 44.9 us: 128-byte pieces of the G table), 16 elsewhere (FP64 at Z = 216
    // with 32 columns spills)
    const int cols = tuning().zpass4_cols;
    pl->C = cols ? cols : (!dbl && Z > 128 ? 32 : 16);
    pl->four = 1;
    return true;
  }
  if (!radix_plan(Z, pl)) return false;
  // columns per workgroup: the zpass_cols knob (4, 8 or 16; 8: 64-byte FP32 /
  // 128-byte FP64 pieces of a z row), halved while a workgroup would hold
  // more than kZpassThreads * kZpassPerThread values or its two buffers
  // would pass 64 KiB of LDS
  const size_t cplx = dbl ? 16 : 8;
  const int maxv = kZpassThreads * kZpassPerThread;
  int C = tuning().zpass_cols;
  while (C > 4 && (Z * C > maxv || (2 * (size_t)Z * C + Z) * cplx > 64 * 1024)) C >>= 1;
  if (Z * C > maxv || (2 * (size_t)Z * C + Z) * cplx > 64 * 1024) return false;
  pl->C = C;
  return true;
}

size_t zpass_lds_bytes(const ZPlan& pl, bool dbl) {
  const size_t cplx = dbl ? 16 : 8;
  if (pl.four) return ((size_t)pl.Z * pl.C + pl.Z) * cplx;
  return (2 * (size_t)pl.Z * pl.C + pl.Z) * cplx;
}

int zpass_blocks(int X, int Y, const ZPlan& pl) {
  const int ncol = (X / 2 + 1) * Y;
  return (ncol + pl.C - 1) / pl.C;
}

template <typename T, int C>
static void zpass_t(const KArgs& a, const ZPlan& pl, int kind, void* spec, const void* table,
                    const void* tw, bool vir, double* partials, unsigned g, size_t lds,
                    hipStream_t s) {
  if (vir)
    k_zpass<T, C, true><<<g, kZpassThreads, lds, s>>>(a, pl, kind, (C2<T>*)spec, (const T*)table,
                                                      (const C2<T>*)tw, partials);
  else
    k_zpass<T, C, false><<<g, kZpassThreads, lds, s>>>(a, pl, kind, (C2<T>*)spec, (const T*)table,
                                                       (const C2<T>*)tw, partials);
}

hipError_t launch_zpass(bool dbl, int kind, int X, int Y, int Z, double bx, double by, double bz,
                        double sigma, const ZPlan& pl, void* spec, const void* table,
                        const void* tw, bool vir, double* partials, hipStream_t s, int zs) {
  KArgs a = make_kargs(X, Y, Z, bx, by, bz, sigma);
  if (zs > 0) {
    if (zs < a.zs) return hipErrorInvalidValue;
    a.zs = zs;
  }
  const unsigned g = (unsigned)zpass_blocks(X, Y, pl);
  const size_t lds = zpass_lds_bytes(pl, dbl);
  if (pl.four) {
    if (Z == 80 && pl.C == 16) {
      if (dbl) zpass4_t<double, 8, 10, 16>(a, kind, spec, table, tw, vir, partials, g, s);
      else zpass4_t<float, 8, 10, 16>(a, kind, spec, table, tw, vir, partials, g, s);
    } else if (Z == 80 && pl.C == 32) {
      if (dbl) zpass4_t<double, 8, 10, 32>(a, kind, spec, table, tw, vir, partials, g, s);
      else zpass4_t<float, 8, 10, 32>(a, kind, spec, table, tw, vir, partials, g, s);
    } else if (Z == 216 && pl.C == 16) {
      if (dbl) zpass4_t<double, 12, 18, 16>(a, kind, spec, table, tw, vir, partials, g, s);
      else zpass4_t<float, 12, 18, 16>(a, kind, spec, table, tw, vir, partials, g, s);
    } else if (Z == 216 && pl.C == 32) {
      if (dbl) zpass4_t<double, 12, 18, 32>(a, kind, spec, table, tw, vir, partials, g, s);
      else zpass4_t<float, 12, 18, 32>(a, kind, spec, table, tw, vir, partials, g, s);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (pl.C != 16 && pl.C != 8 && pl.C != 4) return hipErrorInvalidValue;
  if (dbl) {
    if (pl.C == 16) zpass_t<double, 16>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
    else if (pl.C == 8) zpass_t<double, 8>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
    else zpass_t<double, 4>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
  } else {
    if (pl.C == 16) zpass_t<float, 16>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
    else if (pl.C == 8) zpass_t<float, 8>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
    else zpass_t<float, 4>(a, pl, kind, spec, table, tw, vir, partials, g, lds, s);
  }
  return hipGetLastError();
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

// Separable periodic convolution passes in gather form (CpuGaussCharge.cpp:
// 448-514, CpuConvolution.cpp:134-148): out[i] = sum_{ii=-R..R} k[ii+R]
// in[i+ii] along the pass axis, summed in ii order (FP32: compensated).
// FP32: Kahan-compensated sum of the products (the summation error is carried
// in c; -ffp-contract=off keeps the other steps rounded as written).  The
// product enters as fma(a, b, -c): one rounding instead of two.
template <typename T>
struct ConvAcc {
  T s = 0, c = 0;
  __device__ __forceinline__ void add(T a, T b) {
    const T y = __builtin_fmaf(a, b, -c);
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ T value() const { return s; }
};
template <>
struct ConvAcc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double a, double b) { s += a * b; }
  __device__ __forceinline__ double value() const { return s; }
};

__device__ __forceinline__ int wrap_small(int q, int M) {
  while (q < 0) q += M;
  while (q >= M) q -= M;
  return q;
}

// x pass: each workgroup takes 4 rows (y, z) at a time, one 64-lane wave per
// row; a row and its periodic halo of R values on each side are staged in LDS
// with one coalesced read, then each lane produces 4 consecutive outputs
// (RC > 0: from a register window) reading its inputs from LDS.
template <typename T, int RC, bool SYM>
__global__ void __launch_bounds__(256)
    k_conv_x(int X, int nrows, int R, const T* __restrict__ taps, const T* __restrict__ in,
             T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char conv_smem[];
  const int nt = 2 * R + 1;
  const int pitch = X + 2 * R;
  T* tap = (T*)conv_smem;
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  T* row = tap + ((nt + 3) & ~3) + wv * pitch;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) tap[i] = taps[i];
  for (int r0 = 4 * blockIdx.x; r0 < nrows; r0 += 4 * gridDim.x) {
    const int r = r0 + wv;
    if (r < nrows) {
      const T* src = in + (size_t)r * X;
      for (int i = ln; i < pitch; i += 64) row[i] = src[wrap_small(i - R, X)];
    }
    __syncthreads();
    if (r < nrows) {
      T* dst = out + (size_t)r * X;
      if constexpr (RC > 0) {
        constexpr int W = 4 + 2 * RC, NT = 2 * RC + 1;
        for (int x0 = 4 * ln; x0 < X; x0 += 256) {
          T win[W];
#pragma unroll
          for (int i = 0; i < W; i++) win[i] = row[min(x0 + i, pitch - 1)];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (x0 + u < X) {
              ConvAcc<T> acc;
              if constexpr (SYM) {
                acc.add(tap[RC], win[u + RC]);
#pragma unroll
                for (int i = 1; i <= RC; i++)
                  acc.add(tap[RC + i], win[u + RC - i] + win[u + RC + i]);
              } else {
#pragma unroll
                for (int t = 0; t < NT; t++) acc.add(tap[t], win[u + t]);
              }
              dst[x0 + u] = acc.value();
            }
          }
        }
      } else {
        for (int x = ln; x < X; x += 64) {
          ConvAcc<T> acc;
          for (int t = 0; t < nt; t++) acc.add(tap[t], row[x + t]);
          dst[x] = acc.value();
        }
      }
    }
    __syncthreads();
  }
}

// x pass for X % 4 == 0, X >= 8, radius RC <= 8: workgroup b takes RW
// consecutive rows, one contiguous span of RW * X elements, copies it to LDS
// in 4-element pieces (each row at pitch X + 16, interior at offset 8), fills
// each row's periodic halo from LDS, then works through items of 4
// consecutive outputs of one row: 5 aligned 4-element LDS reads give the
// item's window and one 4-element store writes it.  Items are spread over
// the whole workgroup, so short rows leave no lanes idle.
template <typename T, int RC, bool SYM>
__global__ void __launch_bounds__(256)
    k_conv_x4(int X, int nrows, int RW, const T* __restrict__ taps, const T* __restrict__ in,
              T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char conv_smem[];
  static_assert(RC <= 8, "halo slots");
  constexpr int NT = 2 * RC + 1;
  T* lds = (T*)conv_smem;
  const int P = X + 16, X4 = X >> 2;
  // e / X4 as a multiply-high: exact for e < 2^32 / X4
  const unsigned mg = 0xFFFFFFFFu / (unsigned)X4 + 1u;
  T tp[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) tp[t] = taps[t];
  const int r0 = blockIdx.x * RW;
  const int rw = min(RW, nrows - r0);
  const int n4 = rw * X4;
  const V4<T>* src = (const V4<T>*)(in + (size_t)r0 * X);
  for (int e = threadIdx.x; e < n4; e += 256) {
    const int row = (int)__umulhi((unsigned)e, mg), q = e - row * X4;
    *(V4<T>*)(lds + row * P + 8 + 4 * q) = src[e];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < rw * 2 * RC; e += 256) {
    const int row = e / (2 * RC), h = e - row * (2 * RC);
    T* L = lds + row * P + 8;
    if (h < RC)
      L[h - RC] = L[X - RC + h];  // x = h - RC  <-  X - RC + h
    else
      L[X + h - RC] = L[h - RC];  // x = X + h - RC  <-  h - RC
  }
  __syncthreads();
  V4<T>* dst = (V4<T>*)(out + (size_t)r0 * X);
  for (int e = threadIdx.x; e < n4; e += 256) {
    const int row = (int)__umulhi((unsigned)e, mg), q = e - row * X4;
    const T* L = lds + row * P + 4 * q;  // window input i (x = 4q - RC + i) is w[8 - RC + i]
    T w[20];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const V4<T> v = *(const V4<T>*)(L + 4 * k);
      w[4 * k] = v.x;
      w[4 * k + 1] = v.y;
      w[4 * k + 2] = v.z;
      w[4 * k + 3] = v.w;
    }
    T res[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      ConvAcc<T> acc;
      constexpr int c0 = 8 - RC;
      if constexpr (SYM) {
        acc.add(tp[RC], w[c0 + u + RC]);
#pragma unroll
        for (int i = 1; i <= RC; i++) acc.add(tp[RC + i], w[c0 + u + RC - i] + w[c0 + u + RC + i]);
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) acc.add(tp[t], w[c0 + u + t]);
      }
      res[u] = acc.value();
    }
    V4<T> o;
    o.x = res[0];
    o.y = res[1];
    o.z = res[2];
    o.w = res[3];
    dst[e] = o;
  }
}

// y or z pass: an axis of length N with element stride S; x (contiguous) is
// tiled by 64 lanes, the axis by kConvTJ outputs, `outer` is the remaining
// axis (stride So).  The (kConvTJ + 2R) x 64 input tile is staged in LDS with
// coalesced row reads, each thread then produces kConvTJ/4 outputs.
constexpr int kConvTJ = 32;
// RC > 0: R == RC known at compile time; each thread then produces 8
// consecutive outputs from a register window of 8 + 2R inputs.
template <typename T, int RC, bool SYM>
__global__ void __launch_bounds__(256)
    k_conv_yz(int X, int N, long long S, long long So, int R, const T* __restrict__ taps,
              const T* __restrict__ in, T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char conv_smem[];
  const int nt = 2 * R + 1;
  T* tap = (T*)conv_smem;
  T* tile = tap + ((nt + 3) & ~3);
  const int xl = threadIdx.x & 63, gq = threadIdx.x >> 6;
  const int x = blockIdx.x * 64 + xl;
  const int j0 = blockIdx.y * kConvTJ;
  const size_t obase = (size_t)blockIdx.z * (size_t)So;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) tap[i] = taps[i];
  const int rows = kConvTJ + 2 * R;
  for (int r = gq; r < rows; r += 4) {
    const int j = wrap_small(j0 - R + r, N);
    tile[r * 64 + xl] = x < X ? in[obase + (size_t)x + (size_t)S * j] : (T)0;
  }
  __syncthreads();
  if (x >= X) return;
  if constexpr (RC > 0) {
    constexpr int W = 8 + 2 * RC, NT = 2 * RC + 1;
    const int jb = gq * 8;  // this thread's outputs j0 + jb .. j0 + jb + 7
    T win[W], tp[NT];
#pragma unroll
    for (int i = 0; i < W; i++) win[i] = tile[(jb + i) * 64 + xl];
#pragma unroll
    for (int t = 0; t < NT; t++) tp[t] = tap[t];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if (j0 + jb + u < N) {
        ConvAcc<T> acc;
        if constexpr (SYM) {
          acc.add(tp[RC], win[u + RC]);
#pragma unroll
          for (int i = 1; i <= RC; i++) acc.add(tp[RC + i], win[u + RC - i] + win[u + RC + i]);
        } else {
#pragma unroll
          for (int t = 0; t < NT; t++) acc.add(tp[t], win[u + t]);
        }
        out[obase + (size_t)x + (size_t)S * (j0 + jb + u)] = acc.value();
      }
    }
  } else {
#pragma unroll 2
    for (int u = 0; u < kConvTJ / 4; u++) {
      const int jj = gq + 4 * u;
      if (j0 + jj >= N) break;
      ConvAcc<T> acc;
      for (int t = 0; t < nt; t++) acc.add(tap[t], tile[(jj + t) * 64 + xl]);
      out[obase + (size_t)x + (size_t)S * (j0 + jj)] = acc.value();
    }
  }
}

// y or z pass as register pencils: lane c owns the column at element offset
// (c % cw) + (c / cw) * cs and produces outputs j0 .. j0 + TJ - 1 along the
// axis (length N, element stride S) from a register window of TJ + 2R inputs
// read straight from HBM; consecutive lanes own consecutive elements, so each
// row of the window is one coalesced read per wave.  No LDS, no barrier.
// The row index is wave-uniform: its periodic wrap is scalar work.
// Workgroups are numbered so that the ntile row tiles of one column block
// run on one XCD (workgroups go to the 8 XCDs round-robin), where the rows
// that neighbouring tiles' windows share can come from that XCD's L2.
template <typename T, int RC, bool SYM, int TJ>
__global__ void __launch_bounds__(256)
    k_conv_pencil(int C, int cw, long long cs, int N, long long S, int ncb, int ntile,
                  const T* __restrict__ taps, const T* __restrict__ in, T* __restrict__ out) {
  const int r = (int)(blockIdx.x >> 3);
  const int cb = (r / ntile) * 8 + (int)(blockIdx.x & 7);
  if (cb >= ncb) return;
  const int c = cb * 256 + threadIdx.x;
  if (c >= C) return;
  constexpr int W = TJ + 2 * RC, NT = 2 * RC + 1;
  const int j0 = (r % ntile) * TJ;
  const size_t base = (size_t)(c % cw) + (size_t)(c / cw) * (size_t)cs;
  T tp[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) tp[t] = taps[t];
  int j = (j0 - RC) % N;
  if (j < 0) j += N;
  const T* src = in + base;
  long long o = S * j;
  T win[W];
#pragma unroll
  for (int i = 0; i < W; i++) {
    win[i] = src[o];
    if (++j == N) {
      j = 0;
      o = 0;
    } else {
      o += S;
    }
  }
  T* dst = out + base + S * j0;
#pragma unroll
  for (int u = 0; u < TJ; u++) {
    if (j0 + u < N) {
      ConvAcc<T> acc;
      if constexpr (SYM) {
        acc.add(tp[RC], win[u + RC]);
#pragma unroll
        for (int i = 1; i <= RC; i++) acc.add(tp[RC + i], win[u + RC - i] + win[u + RC + i]);
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) acc.add(tp[t], win[u + t]);
      }
      *dst = acc.value();
    }
    dst += S;
  }
}

// ---------------------------------------------------------------------------
static unsigned grid_for(size_t count, unsigned threads, unsigned cap) {
  size_t g = (count + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

int fused_plane_stride(int X, int Y) { return ((X / 2 + 1) * Y + 31) / 32 * 32; }

hipError_t launch_greens_table(bool dbl, int kind, int X, int Y, int Z, double bx, double by,
                               double bz, double sigma, void* table, hipStream_t s, int zs) {
  KArgs a = make_kargs(X, Y, Z, bx, by, bz, sigma);
  if (zs > 0) {
    if (zs < a.zs) return hipErrorInvalidValue;
    a.zs = zs;
  }
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

hipError_t launch_kpass_finalize(const double* partials, int nblocks, int ncomp, double pref,
                                 double* energy, double* virial, hipStream_t s) {
  k_kpass_finalize<<<1, 256, 0, s>>>(partials, nblocks, ncomp, pref, energy, virial);
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

template <typename T>
__global__ void k_scale_into(T* __restrict__ out, const T* __restrict__ in, size_t n, double f) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i] * (T)f;
}

template <typename T>
__global__ void k_add_into(T* __restrict__ out, const T* __restrict__ in, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = out[i] + in[i];
}

hipError_t launch_grid_scale(bool dbl, void* out, const void* in, size_t n, double f,
                             hipStream_t s) {
  const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  if (dbl)
    k_scale_into<double><<<g > 0 ? g : 1, 256, 0, s>>>((double*)out, (const double*)in, n, f);
  else
    k_scale_into<float><<<g > 0 ? g : 1, 256, 0, s>>>((float*)out, (const float*)in, n, f);
  return hipGetLastError();
}

hipError_t launch_grid_add(bool dbl, void* out, const void* in, size_t n, hipStream_t s) {
  const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  if (dbl)
    k_add_into<double><<<g > 0 ? g : 1, 256, 0, s>>>((double*)out, (const double*)in, n);
  else
    k_add_into<float><<<g > 0 ? g : 1, 256, 0, s>>>((float*)out, (const float*)in, n);
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
                            const void* in, void* out, hipStream_t s, bool sym) {
  const size_t ts = dbl ? 8 : 4;
  const int nt = 2 * R + 1;
  if (pass == 0) {
    const int nrows = Y * Z;
    const size_t vb = 4 * ts;
    if (R == 5 && X % 4 == 0 && X >= 8 && (uintptr_t)in % vb == 0 && (uintptr_t)out % vb == 0 &&
        (size_t)X * nrows < (1ull << 31)) {
      // about 512 four-output items per workgroup
      const int RW = std::max(1, 512 / (X / 4));
      const size_t lds = ts * (size_t)RW * (X + 16);
      if (lds <= 64 * 1024) {
        const unsigned g = (unsigned)((nrows + RW - 1) / RW);
#define GSE_CONVX4(TT, SY) \
  k_conv_x4<TT, 5, SY><<<g, 256, lds, s>>>(X, nrows, RW, (const TT*)taps, (const TT*)in, (TT*)out)
        if (dbl)
          GSE_CONVX4(double, false);
        else if (sym)
          GSE_CONVX4(float, true);
        else
          GSE_CONVX4(float, false);
#undef GSE_CONVX4
        return hipGetLastError();
      }
    }
    const size_t lds = ts * (size_t)(((nt + 3) & ~3) + 4 * (X + 2 * R));
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    const unsigned g = (unsigned)std::min((nrows + 3) / 4, 8192);
    if (dbl) {
      if (R == 5)
        k_conv_x<double, 5, false><<<g, 256, lds, s>>>(X, nrows, R, (const double*)taps,
                                                       (const double*)in, (double*)out);
      else
        k_conv_x<double, 0, false><<<g, 256, lds, s>>>(X, nrows, R, (const double*)taps,
                                                       (const double*)in, (double*)out);
    } else {
      if (R == 5 && sym)
        k_conv_x<float, 5, true><<<g, 256, lds, s>>>(X, nrows, R, (const float*)taps,
                                                     (const float*)in, (float*)out);
      else if (R == 5)
        k_conv_x<float, 5, false><<<g, 256, lds, s>>>(X, nrows, R, (const float*)taps,
                                                      (const float*)in, (float*)out);
      else
        k_conv_x<float, 0, false><<<g, 256, lds, s>>>(X, nrows, R, (const float*)taps,
                                                      (const float*)in, (float*)out);
    }
    return hipGetLastError();
  }
  const int N = pass == 1 ? Y : Z;
  const long long S = pass == 1 ? (long long)X : (long long)X * Y;
  if (R == 5 && (long long)X * Y * Z < (1ll << 31)) {
    // pencils: y pass columns are (x, z) with x fastest, z pass columns the
    // whole xy plane; tiles of 32 outputs when that still gives >= 8 waves
    // per CU on 256 CUs, else 16
    const int C = pass == 1 ? X * Z : X * Y;
    const int cw = pass == 1 ? X : X * Y;
    const long long cs = pass == 1 ? (long long)X * Y : 0;
    const unsigned gx = (unsigned)((C + 255) / 256);
    const bool big = (long long)gx * 4 * ((N + 31) / 32) >= 256 * 8;
    const int ntile = big ? (N + 31) / 32 : (N + 15) / 16;
    const unsigned grid = ((gx + 7) / 8) * 8 * (unsigned)ntile;
#define GSE_PENCIL(TT, SY, TJ)                                                                 \
  k_conv_pencil<TT, 5, SY, TJ><<<grid, 256, 0, s>>>(C, cw, cs, N, S, (int)gx, ntile,       \
                                                     (const TT*)taps, (const TT*)in, (TT*)out)
    if (dbl) {
      if (big) GSE_PENCIL(double, false, 32); else GSE_PENCIL(double, false, 16);
    } else if (sym) {
      if (big) GSE_PENCIL(float, true, 32); else GSE_PENCIL(float, true, 16);
    } else {
      if (big) GSE_PENCIL(float, false, 32); else GSE_PENCIL(float, false, 16);
    }
#undef GSE_PENCIL
    return hipGetLastError();
  }
  const int outer = pass == 1 ? Z : Y;
  const long long So = pass == 1 ? (long long)X * Y : (long long)X;
  const size_t lds = ts * (size_t)(((nt + 3) & ~3) + (kConvTJ + 2 * R) * 64);
  if (lds > 64 * 1024 || outer > 65535) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)((X + 63) / 64), (unsigned)((N + kConvTJ - 1) / kConvTJ),
                  (unsigned)outer);
  if (dbl) {
    if (R == 5)
      k_conv_yz<double, 5, false><<<grid, 256, lds, s>>>(X, N, S, So, R, (const double*)taps,
                                                         (const double*)in, (double*)out);
    else
      k_conv_yz<double, 0, false><<<grid, 256, lds, s>>>(X, N, S, So, R, (const double*)taps,
                                                         (const double*)in, (double*)out);
  } else {
    if (R == 5 && sym)
      k_conv_yz<float, 5, true><<<grid, 256, lds, s>>>(X, N, S, So, R, (const float*)taps,
                                                       (const float*)in, (float*)out);
    else if (R == 5)
      k_conv_yz<float, 5, false><<<grid, 256, lds, s>>>(X, N, S, So, R, (const float*)taps,
                                                        (const float*)in, (float*)out);
    else
      k_conv_yz<float, 0, false><<<grid, 256, lds, s>>>(X, N, S, So, R, (const float*)taps,
                                                        (const float*)in, (float*)out);
  }
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
