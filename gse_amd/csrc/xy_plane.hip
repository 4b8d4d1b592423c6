// x-y plane transforms of the fused FFT solve (CpuFFTSolver.cpp:75-95 does
// the whole solve as one 3D r2c, the G(k) multiply and one 3D c2r; the fused
// form runs the x-y part of both transforms here and the z part in the z
// pass).  One workgroup per z plane, everything in LDS:
//
// r2c: rows y and y + ceil(Y/2) are packed as one complex row
// z = rho[y] + i rho[y + Yh] and transformed along x (length X, mixed-radix
// Stockham); the two real spectra are separated with
//   A[k] = (Z[k] + conj Z[-k]) / 2,  B[k] = (Z[k] - conj Z[-k]) / (2i),
// k <= X/2, into a [y][kx] plane, which is transformed along y and stored
// as spec[z][ky][kx].
// c2r: the reverse.  The inverse y transform gives the row spectra; each
// pair is extended to all X by Hermitian symmetry (the imaginary part of
// the k = 0 and k = X/2 terms ignored, as a c2r transform does), packed as
// A + iB, transformed back along x, and the real and imaginary parts are the
// two rows.
#include <algorithm>

#include "gse_device.h"
#include "gse_internal.h"
#include "zdft_device.h"

namespace gse {

constexpr int kPlaneThreads = 1024;
constexpr int kPlaneLoads = 8;  // global loads in flight per thread
constexpr size_t kPlaneLdsMax = 160 * 1024;

// One Stockham stage of radix P over ncol columns of length L held in LDS
// with element (position i, column c) at i * si + c * sc; cfast: consecutive
// threads take consecutive columns (else consecutive positions).
template <typename T, int P>
__device__ __forceinline__ void plane_stage(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst,
                                            const C2<T>* __restrict__ W, int L, int Tsub,
                                            int ncol, int si, int sc, bool cfast, bool inv) {
  const int m = L / P;
  const int wstep = L / (P * Tsub);
  C2<T> root[P];
#pragma unroll
  for (int r = 0; r < P; r++) {
    root[r] = W[r * m];
    if (inv) root[r].y = -root[r].y;
  }
  const int n = m * ncol;
  for (int w = threadIdx.x; w < n; w += blockDim.x) {
    int i, c;
    if (cfast) {
      i = w / ncol;
      c = w - i * ncol;
    } else {
      c = w / m;
      i = w - c * m;
    }
    const int k = i % Tsub;
    C2<T> x[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
      C2<T> v = src[(i + q * m) * si + c * sc];
      if (q > 0 && k > 0) v = zc_mul(v, W[q * k * wstep], inv);
      x[q] = v;
    }
    const int j = (i - k) * P + k;
    if (zdft_small<T, P>(x, inv)) {
#pragma unroll
      for (int s = 0; s < P; s++) dst[(j + s * Tsub) * si + c * sc] = x[s];
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
        dst[(j + s * Tsub) * si + c * sc] = C2<T>{yr, yi};
      }
    }
  }
}

// the whole transform (a -> b -> a ...); returns the buffer holding the
// result.  Starts and ends with a barrier-separated state.
template <typename T>
__device__ C2<T>* plane_fft(C2<T>* a, C2<T>* b, const C2<T>* W, const ZPlan& pl, int ncol, int si,
                            int sc, bool cfast, bool inv) {
  int Tsub = 1;
  for (int st = 0; st < pl.nst; st++) {
    const int p = pl.p[st];
    switch (p) {
      case 2: plane_stage<T, 2>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      case 3: plane_stage<T, 3>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      case 4: plane_stage<T, 4>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      case 5: plane_stage<T, 5>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      case 6: plane_stage<T, 6>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      case 7: plane_stage<T, 7>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
      default: plane_stage<T, 8>(a, b, W, pl.Z, Tsub, ncol, si, sc, cfast, inv); break;
    }
    __syncthreads();
    C2<T>* t = a;
    a = b;
    b = t;
    Tsub *= p;
  }
  return a;
}

__host__ __device__ inline int plane_buf_elems(int X, int Y) {
  const int Xh = X / 2 + 1, Yh = (Y + 1) / 2;
  return X * Yh > Xh * Y ? X * Yh : Xh * Y;
}

static size_t plane_lds_bytes(int X, int Y, bool dbl) {
  return (2 * (size_t)plane_buf_elems(X, Y) + X + Y) * (dbl ? 16 : 8);
}

template <typename T>
__global__ void __launch_bounds__(kPlaneThreads)
    k_xy_r2c(const T* __restrict__ rho, C2<T>* __restrict__ spec, int X, int Y, int zs, ZPlan px,
             ZPlan py, const C2<T>* __restrict__ twx, const C2<T>* __restrict__ twy) {
  extern __shared__ __align__(16) char psm[];
  const int Xh = X / 2 + 1, Yh = (Y + 1) / 2;
  const int nbuf = plane_buf_elems(X, Y);
  C2<T>* A = (C2<T>*)psm;
  C2<T>* B = A + nbuf;
  C2<T>* Wx = B + nbuf;
  C2<T>* Wy = Wx + X;
  const size_t plane = (size_t)X * Y;
  const T* src = rho + (size_t)blockIdx.x * plane;
  // packed rows: A[rp X + x] = rho[rp][x] + i rho[rp + Yh][x]; every load
  // of a chunk is issued before its LDS stores
  const int npk = X * Yh;
  for (int base = 0; base < npk; base += kPlaneLoads * kPlaneThreads) {
    T a[kPlaneLoads], b[kPlaneLoads];
#pragma unroll
    for (int u = 0; u < kPlaneLoads; u++) {
      const int idx = base + u * kPlaneThreads + threadIdx.x;
      const int rp = idx / X, x = idx - rp * X;
      a[u] = idx < npk ? src[(size_t)rp * X + x] : (T)0;
      b[u] = idx < npk && rp + Yh < Y ? src[(size_t)(rp + Yh) * X + x] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kPlaneLoads; u++) {
      const int idx = base + u * kPlaneThreads + threadIdx.x;
      if (idx < npk) A[idx] = C2<T>{a[u], b[u]};
    }
  }
  for (int i = threadIdx.x; i < X; i += blockDim.x) Wx[i] = twx[i];
  for (int i = threadIdx.x; i < Y; i += blockDim.x) Wy[i] = twy[i];
  __syncthreads();
  C2<T>* R = plane_fft<T>(A, B, Wx, px, Yh, 1, X, false, false);
  C2<T>* O = R == A ? B : A;
  // separate the two real rows' spectra into O[y Xh + kx], kx <= X/2
  for (int idx = threadIdx.x; idx < Xh * Yh; idx += blockDim.x) {
    const int rp = idx / Xh, k = idx - rp * Xh;
    const C2<T> zk = R[rp * X + k];
    const C2<T> zm = R[rp * X + (k == 0 ? 0 : X - k)];
    O[rp * Xh + k] = C2<T>{(T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y)};
    if (rp + Yh < Y)
      O[(rp + Yh) * Xh + k] = C2<T>{(T)0.5 * (zk.y + zm.y), (T)0.5 * (zm.x - zk.x)};
  }
  __syncthreads();
  C2<T>* F = plane_fft<T>(O, R, Wy, py, Xh, Xh, 1, true, false);
  C2<T>* dst = spec + (size_t)blockIdx.x * zs;
  for (int idx = threadIdx.x; idx < Xh * Y; idx += blockDim.x) dst[idx] = F[idx];
}

template <typename T>
__global__ void __launch_bounds__(kPlaneThreads)
    k_xy_c2r(const C2<T>* __restrict__ spec, T* __restrict__ phi, int X, int Y, int zs, ZPlan px,
             ZPlan py, const C2<T>* __restrict__ twx, const C2<T>* __restrict__ twy) {
  extern __shared__ __align__(16) char psm[];
  const int Xh = X / 2 + 1, Yh = (Y + 1) / 2;
  const int nbuf = plane_buf_elems(X, Y);
  C2<T>* A = (C2<T>*)psm;
  C2<T>* B = A + nbuf;
  C2<T>* Wx = B + nbuf;
  C2<T>* Wy = Wx + X;
  const C2<T>* src = spec + (size_t)blockIdx.x * zs;
  const int nsp = Xh * Y;
  for (int base = 0; base < nsp; base += kPlaneLoads * kPlaneThreads) {
    C2<T> v[kPlaneLoads];
#pragma unroll
    for (int u = 0; u < kPlaneLoads; u++) {
      const int idx = base + u * kPlaneThreads + threadIdx.x;
      v[u] = idx < nsp ? src[idx] : C2<T>{(T)0, (T)0};
    }
#pragma unroll
    for (int u = 0; u < kPlaneLoads; u++) {
      const int idx = base + u * kPlaneThreads + threadIdx.x;
      if (idx < nsp) A[idx] = v[u];
    }
  }
  for (int i = threadIdx.x; i < X; i += blockDim.x) Wx[i] = twx[i];
  for (int i = threadIdx.x; i < Y; i += blockDim.x) Wy[i] = twy[i];
  __syncthreads();
  C2<T>* R = plane_fft<T>(A, B, Wy, py, Xh, Xh, 1, true, true);
  C2<T>* O = R == A ? B : A;
  // rows rp and rp + Yh: Z[k] = A[k] + i B[k] over all k < X, A and B
  // extended by Hermitian symmetry
  const bool even = (X % 2) == 0;
  for (int idx = threadIdx.x; idx < X * Yh; idx += blockDim.x) {
    const int rp = idx / X, k = idx - rp * X;
    const bool lo = k < Xh;
    const int kk = lo ? k : X - k;
    const bool real_only = kk == 0 || (even && 2 * kk == X);
    C2<T> a = R[rp * Xh + kk];
    C2<T> b = rp + Yh < Y ? R[(rp + Yh) * Xh + kk] : C2<T>{(T)0, (T)0};
    if (real_only) {
      a.y = (T)0;
      b.y = (T)0;
    } else if (!lo) {
      a.y = -a.y;
      b.y = -b.y;
    }
    O[idx] = C2<T>{a.x - b.y, a.y + b.x};
  }
  __syncthreads();
  C2<T>* F = plane_fft<T>(O, R, Wx, px, Yh, 1, X, false, true);
  T* dst = phi + (size_t)blockIdx.x * X * Y;
  for (int idx = threadIdx.x; idx < X * Yh; idx += blockDim.x) {
    const int rp = idx / X, x = idx - rp * X;
    const C2<T> v = F[idx];
    dst[(size_t)rp * X + x] = v.x;
    if (rp + Yh < Y) dst[(size_t)(rp + Yh) * X + x] = v.y;
  }
}

bool xy_plane_ok(int X, int Y, bool dbl) {
  ZPlan px, py;
  return X >= 2 && Y >= 2 && radix_plan(X, &px) && radix_plan(Y, &py) &&
         plane_lds_bytes(X, Y, dbl) <= kPlaneLdsMax;
}

hipError_t launch_xy_r2c(bool dbl, int X, int Y, int Z, const void* rho, void* spec,
                         const void* twx, const void* twy, hipStream_t s, int zs) {
  ZPlan px, py;
  if (Z < 1 || !xy_plane_ok(X, Y, dbl) || !radix_plan(X, &px) || !radix_plan(Y, &py))
    return hipErrorInvalidValue;
  if (zs == 0) zs = (X / 2 + 1) * Y;
  if (zs < (X / 2 + 1) * Y) return hipErrorInvalidValue;
  const size_t lds = plane_lds_bytes(X, Y, dbl);
  if (dbl)
    k_xy_r2c<double><<<Z, kPlaneThreads, lds, s>>>((const double*)rho, (C2<double>*)spec, X, Y,
                                                   zs, px, py, (const C2<double>*)twx,
                                                   (const C2<double>*)twy);
  else
    k_xy_r2c<float><<<Z, kPlaneThreads, lds, s>>>((const float*)rho, (C2<float>*)spec, X, Y, zs,
                                                  px, py, (const C2<float>*)twx,
                                                  (const C2<float>*)twy);
  return hipGetLastError();
}

hipError_t launch_xy_c2r(bool dbl, int X, int Y, int Z, const void* spec, void* phi,
                         const void* twx, const void* twy, hipStream_t s, int zs) {
  ZPlan px, py;
  if (Z < 1 || !xy_plane_ok(X, Y, dbl) || !radix_plan(X, &px) || !radix_plan(Y, &py))
    return hipErrorInvalidValue;
  if (zs == 0) zs = (X / 2 + 1) * Y;
  if (zs < (X / 2 + 1) * Y) return hipErrorInvalidValue;
  const size_t lds = plane_lds_bytes(X, Y, dbl);
  if (dbl)
    k_xy_c2r<double><<<Z, kPlaneThreads, lds, s>>>((const C2<double>*)spec, (double*)phi, X, Y,
                                                   zs, px, py, (const C2<double>*)twx,
                                                   (const C2<double>*)twy);
  else
    k_xy_c2r<float><<<Z, kPlaneThreads, lds, s>>>((const C2<float>*)spec, (float*)phi, X, Y, zs,
                                                  px, py, (const C2<float>*)twx,
                                                  (const C2<float>*)twy);
  return hipGetLastError();
}

}  // namespace gse
