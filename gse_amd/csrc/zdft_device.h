// Complex values and small in-register DFTs shared by the z pass and the
// x-y plane transforms.  Sign convention: forward w = exp(-2 pi i / P),
// inverse w = exp(+2 pi i / P), both unnormalised.
#pragma once

#include <hip/hip_runtime.h>

namespace gse {

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};

// Small DFTs y_s = sum_q x_q w^(q s), w = exp(-2 pi i / P) (exp(+2 pi i / P)
// for the inverse), in place, as radix-2 FFT steps for P = 2, 4, 8 and the
// three-point split for P = 3; other radices use the direct sum.
template <typename T>
__device__ __forceinline__ C2<T> zc_add(C2<T> a, C2<T> b) {
  return C2<T>{a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ C2<T> zc_sub(C2<T> a, C2<T> b) {
  return C2<T>{a.x - b.x, a.y - b.y};
}
// z * (-i) forward, z * (+i) inverse
template <typename T>
__device__ __forceinline__ C2<T> zc_rot(C2<T> z, bool inv) {
  return inv ? C2<T>{-z.y, z.x} : C2<T>{z.y, -z.x};
}
template <typename T>
__device__ __forceinline__ void zdft4(C2<T>& x0, C2<T>& x1, C2<T>& x2, C2<T>& x3, bool inv) {
  const C2<T> a0 = zc_add(x0, x2), a1 = zc_sub(x0, x2);
  const C2<T> a2 = zc_add(x1, x3), a3 = zc_rot(zc_sub(x1, x3), inv);
  x0 = zc_add(a0, a2);
  x2 = zc_sub(a0, a2);
  x1 = zc_add(a1, a3);
  x3 = zc_sub(a1, a3);
}
template <typename T, int P>
__device__ __forceinline__ bool zdft_small(C2<T> (&x)[P], bool inv) {
  if constexpr (P == 2) {
    const C2<T> a = x[0];
    x[0] = zc_add(a, x[1]);
    x[1] = zc_sub(a, x[1]);
    return true;
  } else if constexpr (P == 3) {
    const T c = (T)0.86602540378443864676;  // sqrt(3) / 2
    const C2<T> s12 = zc_add(x[1], x[2]), d = zc_sub(x[1], x[2]);
    const C2<T> t{x[0].x - (T)0.5 * s12.x, x[0].y - (T)0.5 * s12.y};
    const C2<T> u = zc_rot(C2<T>{c * d.x, c * d.y}, inv);
    x[0] = zc_add(x[0], s12);
    x[1] = zc_add(t, u);
    x[2] = zc_sub(t, u);
    return true;
  } else if constexpr (P == 4) {
    zdft4(x[0], x[1], x[2], x[3], inv);
    return true;
  } else if constexpr (P == 8) {
    const T r = (T)0.70710678118654752440;  // sqrt(1/2)
    C2<T> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    C2<T> o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    zdft4(e0, e1, e2, e3, inv);
    zdft4(o0, o1, o2, o3, inv);
    // o_k w8^k, w8 = (1 -+ i) / sqrt(2)
    const C2<T> t1 = inv ? C2<T>{r * (o1.x - o1.y), r * (o1.x + o1.y)}
                         : C2<T>{r * (o1.x + o1.y), r * (o1.y - o1.x)};
    const C2<T> t2 = zc_rot(o2, inv);
    const C2<T> t3 = inv ? C2<T>{-r * (o3.x + o3.y), r * (o3.x - o3.y)}
                         : C2<T>{r * (o3.y - o3.x), -r * (o3.x + o3.y)};
    x[0] = zc_add(e0, o0);
    x[4] = zc_sub(e0, o0);
    x[1] = zc_add(e1, t1);
    x[5] = zc_sub(e1, t1);
    x[2] = zc_add(e2, t2);
    x[6] = zc_sub(e2, t2);
    x[3] = zc_add(e3, t3);
    x[7] = zc_sub(e3, t3);
    return true;
  } else {
    return false;
  }
}

// In-register DFTs of compile-time length L: x[0..L) in place, twiddles
// w_L^j = W[j ZW / L] from a table W[j] = exp(-2 pi i j / ZW) (L | ZW).  A
// length L = P Q splits by decimation in time,
//   X[q + Q s] = sum_p w_P^(p s) w_L^(p q) DFT_Q(x[p + P m])[q],
// down to the radix-2/3/4/8 butterflies above and direct sums for other
// prime lengths.
template <int L>
constexpr int reg_dft_radix() {
  return (L % 4 == 0 && L != 4 && L != 8) ? 4
         : (L % 2 == 0 && L != 2 && L != 4 && L != 8) ? 2
         : (L % 3 == 0 && L != 3) ? 3
         : (L % 5 == 0 && L != 5) ? 5
         : (L % 7 == 0 && L != 7) ? 7
                                  : L;
}

template <typename T>
__device__ __forceinline__ C2<T> zc_mul(C2<T> a, C2<T> w, bool inv) {
  if (inv) w.y = -w.y;
  return C2<T>{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

// x[0..L) in place; ZW: the table length (W[j] = exp(-2 pi i j / ZW)), L | ZW
template <typename T, int L, int ZW>
__device__ __forceinline__ void reg_dft(C2<T>* x, const C2<T>* __restrict__ W, bool inv) {
  if constexpr (L == 1) {
    return;
  } else if constexpr (L == 2 || L == 3 || L == 4 || L == 8) {
    C2<T> y[L];
#pragma unroll
    for (int i = 0; i < L; i++) y[i] = x[i];
    zdft_small<T, L>(y, inv);
#pragma unroll
    for (int i = 0; i < L; i++) x[i] = y[i];
  } else if constexpr (reg_dft_radix<L>() == L) {  // 5, 7: direct sums
    C2<T> y[L];
#pragma unroll
    for (int s2 = 0; s2 < L; s2++) {
      C2<T> acc = x[0];
#pragma unroll
      for (int q = 1; q < L; q++) {
        const C2<T> t = zc_mul(x[q], W[((q * s2) % L) * (ZW / L)], inv);
        acc = zc_add(acc, t);
      }
      y[s2] = acc;
    }
#pragma unroll
    for (int i = 0; i < L; i++) x[i] = y[i];
  } else {
    constexpr int P = reg_dft_radix<L>(), Q = L / P;
    C2<T> y[P][Q];
#pragma unroll
    for (int p2 = 0; p2 < P; p2++) {
#pragma unroll
      for (int m = 0; m < Q; m++) y[p2][m] = x[p2 + P * m];
      reg_dft<T, Q, ZW>(y[p2], W, inv);
#pragma unroll
      for (int q = 1; q < Q; q++)
        if (p2 > 0) y[p2][q] = zc_mul(y[p2][q], W[p2 * q * (ZW / L)], inv);
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      C2<T> z[P];
#pragma unroll
      for (int p2 = 0; p2 < P; p2++) z[p2] = y[p2][q];
      reg_dft<T, P, ZW>(z, W, inv);
#pragma unroll
      for (int s2 = 0; s2 < P; s2++) x[q + Q * s2] = z[s2];
    }
  }
}


}  // namespace gse
