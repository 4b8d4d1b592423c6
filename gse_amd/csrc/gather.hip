// Grid -> particle gather: the reference force
// (CpuGaussCharge::interpolateLoop, CpuGaussCharge.cpp:105-172: sphere
// support r2 < rcut2 inside the (2n+1)^3 cube) and the build-defined
// cube-support outputs (exact gradient, per-particle potential, virial
// moments) in one pass over the cube.
//
// One workgroup per bin (the binning of the spread).  The bin's potential
// brick of (B+2n)^3 values is staged once into LDS with wrapped, x-coalesced
// loads; each wave then takes one particle at a time: lanes compute the 3(2n+1)
// one-dimensional offsets and Gaussian factors, lanes own (tx,ty) columns of
// the cube and walk tz with the z factors held wave-uniform, and the partial
// sums are reduced across the wave with a fixed butterfly.
#include <math.h>

#include "gse_device.h"
#include "gse_internal.h"

namespace gse {

template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
  const int chunk = (nwg + 7) >> 3;
  return (wg & 7) * chunk + (wg >> 3);
}

constexpr int kMaxK = 5;

struct GatherArgs {
  Geom g;
  BinShape b;
  double inv2s2;
  double rc2;
  double prefF;  // (2 pi)^-3/2 h^3 sigma^-5   (CpuGaussCharge.cpp:128)
  double prefP;  // (2 pi sigma^2)^-3/2 h^3
};

// FL: GatherFlags bit mask; NZC: 2nz+1 known at compile time (0 = runtime)
template <typename CT, typename AT, int NZC, int FL>
__global__ void __launch_bounds__(256)
    k_gather(GatherArgs a, const V4<CT>* __restrict__ sorted, const int* __restrict__ perm,
             const int* __restrict__ starts, const CT* __restrict__ phi, GatherOut o) {
  extern __shared__ char smem_raw[];
  CT* brick = (CT*)smem_raw;
  const Geom& g = a.g;
  const BinShape& b = a.b;
  const int bin = xcd_remap(blockIdx.x, gridDim.x);
  if (bin >= b.nbins) return;
  const int s0 = starts[bin], s1 = starts[bin + 1];
  if (s0 == s1) return;
  const int bxi = bin % b.nbx, byi = (bin / b.nbx) % b.nby, bzi = bin / (b.nbx * b.nby);
  const int Lx = b.Lx, Ly = b.Ly, LxLy = b.Lx * b.Ly;
  const int vol = LxLy * b.Lz;
  {
    const int gx0 = bxi * b.Bx, gy0 = byi * b.By, gz0 = bzi * b.Bz;
    for (int j = threadIdx.x; j < vol; j += blockDim.x) {
      const int lx = j % Lx;
      const int t = j / Lx;
      const int ly = t % Ly;
      const int lz = t / Ly;
      int gx = gx0 + lx, gy = gy0 + ly, gz = gz0 + lz;
      while (gx >= g.X) gx -= g.X;
      while (gy >= g.Y) gy -= g.Y;
      while (gz >= g.Z) gz -= g.Z;
      brick[j] = phi[gx + (size_t)g.X * (gy + (size_t)g.Y * gz)];
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1;
  const int NZ = NZC > 0 ? NZC : 2 * g.nz + 1;
  const int NXY = NX * NY;
  const int KI = (NXY + 63) >> 6;
  const CT rc2 = (CT)a.rc2;
  const CT inv2s2 = (CT)a.inv2s2;

  // lane roles for the 1D factors
  int axis, jj;
  if (lane < NX) {
    axis = 0; jj = lane;
  } else if (lane < NX + NY) {
    axis = 1; jj = lane - NX;
  } else if (lane < NX + NY + NZ) {
    axis = 2; jj = lane - NX - NY;
  } else {
    axis = -1; jj = 0;
  }
  const double hsel = axis == 0 ? g.hx : axis == 1 ? g.hy : g.hz;
  const double ihsel = axis == 0 ? g.ihx : axis == 1 ? g.ihy : g.ihz;
  const int nsel = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;

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

  constexpr bool SPH = (FL & GF_SPHERE) != 0;
  constexpr bool CUB = (FL & GF_CUBE) != 0;
  constexpr bool POT = (FL & GF_POT) != 0;
  constexpr bool VIR = (FL & GF_VIRIAL) != 0;
  constexpr bool NEED_B0 = CUB || POT || VIR;
  constexpr bool NEED_B1 = CUB || VIR;
  constexpr int NZA = NZC > 0 ? NZC : 1;

  for (int i = s0 + wid; i < s1; i += nw) {
    const V4<CT> p = sorted[i];
    const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
    const double ps = axis == 0 ? x : axis == 1 ? y : z;
    // centre cell (int)round(x*inv_h) (CpuGaussCharge.cpp:139-141); this
    // lane's grid index t = c - n + j and offset d = x - t*h (:146-155)
    const int c = (int)round(ps * ihsel);
    const double dd = sub_rn(ps, mul_rn((double)(c - nsel + jj), hsel));
    const CT d = (CT)dd;
    const CT w = axis >= 0 ? gexp(mul_rn(mul_rn(-d, d), inv2s2)) : (CT)0;

    const int cxo = (int)round(x * g.ihx) - g.nx;
    const int cyo = (int)round(y * g.ihy) - g.ny;
    const int czo = (int)round(z * g.ihz) - g.nz;
    const int ox = wrapi(cxo, g.X) - bxi * b.Bx;
    const int oy = wrapi(cyo, g.Y) - byi * b.By;
    const int oz = wrapi(czo, g.Z) - bzi * b.Bz;
    const int base = ox + Lx * (oy + Ly * oz);

    CT wzr[NZA], dzr[NZA];
    if (NZC > 0) {
#pragma unroll
      for (int t = 0; t < NZA; t++) {
        wzr[t] = readlane(w, NX + NY + t);
        dzr[t] = readlane(d, NX + NY + t);
      }
    }

    AT fsx = 0, fsy = 0, fsz = 0, fcx = 0, fcy = 0, fcz = 0, pp = 0;
    AT vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
#pragma unroll
    for (int k = 0; k < kMaxK; k++) {
      if (k >= KI) break;
      const CT wx = shfl(w, k_tx[k]);
      const CT dx = shfl(d, k_tx[k]);
      const CT wy = shfl(w, NX + k_ty[k]);
      const CT dy = shfl(d, NX + k_ty[k]);
      const CT dxy2 = add_rn(mul_rn(dx, dx), mul_rn(dy, dy));
      const int a0 = base + k_off[k];
      AT B0 = 0, B1 = 0, B2 = 0, A0 = 0, A1 = 0;
      if (k_ok[k]) {
        if (NZC > 0) {
#pragma unroll
          for (int t = 0; t < NZA; t++) {
            const CT ph = brick[a0 + t * LxLy];
            const AT wph = (AT)(wzr[t] * ph);
            const AT dzt = (AT)dzr[t];
            if (NEED_B0) B0 += wph;
            if (NEED_B1) B1 += wph * dzt;
            if (VIR) B2 += wph * dzt * dzt;
            if (SPH) {
              const CT r2 = add_rn(dxy2, mul_rn(dzr[t], dzr[t]));
              if (r2 < rc2) {
                A0 += wph;
                A1 += wph * dzt;
              }
            }
          }
        } else {
          for (int t = 0; t < NZ; t++) {
            const CT wzt = readlane(w, NX + NY + t);
            const CT dzc = readlane(d, NX + NY + t);
            const CT ph = brick[a0 + t * LxLy];
            const AT wph = (AT)(wzt * ph);
            const AT dzt = (AT)dzc;
            if (NEED_B0) B0 += wph;
            if (NEED_B1) B1 += wph * dzt;
            if (VIR) B2 += wph * dzt * dzt;
            if (SPH) {
              const CT r2 = add_rn(dxy2, mul_rn(dzc, dzc));
              if (r2 < rc2) {
                A0 += wph;
                A1 += wph * dzt;
              }
            }
          }
        }
        const AT wxy = (AT)wx * (AT)wy;
        const AT ax = (AT)dx, ay = (AT)dy;
        if (SPH) {
          fsx += wxy * ax * A0;
          fsy += wxy * ay * A0;
          fsz += wxy * A1;
        }
        if (CUB) {
          fcx += wxy * ax * B0;
          fcy += wxy * ay * B0;
          fcz += wxy * B1;
        }
        if (POT) pp += wxy * B0;
        if (VIR) {
          vxx += wxy * ax * ax * B0;
          vxy += wxy * ax * ay * B0;
          vxz += wxy * ax * B1;
          vyy += wxy * ay * ay * B0;
          vyz += wxy * ay * B1;
          vzz += wxy * B2;
        }
      }
    }
    if (SPH) {
      fsx = wave_sum(fsx);
      fsy = wave_sum(fsy);
      fsz = wave_sum(fsz);
    }
    if (CUB) {
      fcx = wave_sum(fcx);
      fcy = wave_sum(fcy);
      fcz = wave_sum(fcz);
    }
    if (POT) pp = wave_sum(pp);
    if (VIR) {
      vxx = wave_sum(vxx);
      vxy = wave_sum(vxy);
      vxz = wave_sum(vxz);
      vyy = wave_sum(vyy);
      vyz = wave_sum(vyz);
      vzz = wave_sum(vzz);
    }
    if (lane == 0) {
      const int oi = perm[i];
      const double ds = o.dscale ? *o.dscale : 1.0;
      const double fq = (double)p.w * a.prefF * o.fscale * ds;
      const double vq = (double)p.w * a.prefF * o.vscale;
      AT* const ofx = (AT*)o.fx;
      AT* const ofy = (AT*)o.fy;
      AT* const ofz = (AT*)o.fz;
      if (SPH && ofx) {
        ofx[oi] = (AT)(fq * (double)fsx);
        ofy[oi] = (AT)(fq * (double)fsy);
        ofz[oi] = (AT)(fq * (double)fsz);
      }
      if (CUB && o.cx) {
        ((AT*)o.cx)[oi] = (AT)(fq * (double)fcx);
        ((AT*)o.cy)[oi] = (AT)(fq * (double)fcy);
        ((AT*)o.cz)[oi] = (AT)(fq * (double)fcz);
      }
      if (POT && o.pot) ((AT*)o.pot)[oi] = (AT)(a.prefP * o.pscale * ds * (double)pp);
      if (VIR && o.vir6) {
        double* v = o.vir6;
        const size_t stride = (size_t)o.vir_stride;
        v[0 * stride + oi] = vq * (double)vxx;
        v[1 * stride + oi] = vq * (double)vxy;
        v[2 * stride + oi] = vq * (double)vxz;
        v[3 * stride + oi] = vq * (double)vyy;
        v[4 * stride + oi] = vq * (double)vyz;
        v[5 * stride + oi] = vq * (double)vzz;
      }
    }
  }
}

// Fallback for supports too wide for the lane layout or the LDS brick: one
// wave per particle, every cube point evaluated with its own factors and the
// potential read from global memory.
template <typename CT, typename AT>
__global__ void __launch_bounds__(256)
    k_gather_global(GatherArgs a, int flags, const V4<CT>* __restrict__ sorted,
                    const int* __restrict__ perm, const int* __restrict__ starts,
                    const CT* __restrict__ phi, GatherOut o) {
  const Geom& g = a.g;
  const int n = starts[a.b.nbins];
  const int lane = threadIdx.x & 63;
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (gridDim.x * blockDim.x) >> 6;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const CT inv2s2 = (CT)a.inv2s2;
  const CT rc2 = (CT)a.rc2;
  for (int i = gw; i < n; i += nwave) {
    const V4<CT> p = sorted[i];
    const double x = (double)p.x, y = (double)p.y, z = (double)p.z;
    const int cx = (int)round(x * g.ihx) - g.nx;
    const int cy = (int)round(y * g.ihy) - g.ny;
    const int cz = (int)round(z * g.ihz) - g.nz;
    AT f[13];
    for (int c = 0; c < 13; c++) f[c] = 0;
    for (int idx = lane; idx < NX * NY * NZ; idx += 64) {
      const int tx = idx % NX, t2 = idx / NX, ty = t2 % NY, tz = t2 / NY;
      const CT dx = (CT)sub_rn(x, mul_rn((double)(cx + tx), g.hx));
      const CT dy = (CT)sub_rn(y, mul_rn((double)(cy + ty), g.hy));
      const CT dz = (CT)sub_rn(z, mul_rn((double)(cz + tz), g.hz));
      const CT r2 = add_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy)), mul_rn(dz, dz));
      const int gx = wrapi(cx + tx, g.X), gy = wrapi(cy + ty, g.Y), gz = wrapi(cz + tz, g.Z);
      const CT ph = phi[gx + (size_t)g.X * (gy + (size_t)g.Y * gz)];
      const CT wt = gexp(mul_rn(mul_rn(-dx, dx), inv2s2)) * gexp(mul_rn(mul_rn(-dy, dy), inv2s2)) *
                    gexp(mul_rn(mul_rn(-dz, dz), inv2s2));
      const AT s = (AT)(wt * ph);
      if (r2 < rc2) {
        f[0] += s * (AT)dx;
        f[1] += s * (AT)dy;
        f[2] += s * (AT)dz;
      }
      f[3] += s * (AT)dx;
      f[4] += s * (AT)dy;
      f[5] += s * (AT)dz;
      f[6] += s;
      f[7] += s * (AT)dx * (AT)dx;
      f[8] += s * (AT)dx * (AT)dy;
      f[9] += s * (AT)dx * (AT)dz;
      f[10] += s * (AT)dy * (AT)dy;
      f[11] += s * (AT)dy * (AT)dz;
      f[12] += s * (AT)dz * (AT)dz;
    }
    for (int c = 0; c < 13; c++) f[c] = wave_sum(f[c]);
    if (lane == 0) {
      const int oi = perm[i];
      const double ds = o.dscale ? *o.dscale : 1.0;
      const double fq = (double)p.w * a.prefF * o.fscale * ds;
      const double vq = (double)p.w * a.prefF * o.vscale;
      if ((flags & GF_SPHERE) && o.fx) {
        ((AT*)o.fx)[oi] = (AT)(fq * (double)f[0]);
        ((AT*)o.fy)[oi] = (AT)(fq * (double)f[1]);
        ((AT*)o.fz)[oi] = (AT)(fq * (double)f[2]);
      }
      if ((flags & GF_CUBE) && o.cx) {
        ((AT*)o.cx)[oi] = (AT)(fq * (double)f[3]);
        ((AT*)o.cy)[oi] = (AT)(fq * (double)f[4]);
        ((AT*)o.cz)[oi] = (AT)(fq * (double)f[5]);
      }
      if ((flags & GF_POT) && o.pot) ((AT*)o.pot)[oi] = (AT)(a.prefP * o.pscale * ds * (double)f[6]);
      if ((flags & GF_VIRIAL) && o.vir6) {
        const size_t st = (size_t)o.vir_stride;
        for (int c = 0; c < 6; c++) o.vir6[c * st + oi] = vq * (double)f[7 + c];
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <typename CT, typename AT, int FL>
static hipError_t gather_nz(const GatherArgs& a, const Binning& w, const void* phi,
                            const GatherOut& o, size_t lds, unsigned nwg, hipStream_t s) {
  const V4<CT>* srt = (const V4<CT>*)w.sorted;
  const CT* ph = (const CT*)phi;
  switch (2 * a.g.nz + 1) {
#define GSE_G_CASE(NZV)                                                                     \
  case NZV:                                                                                 \
    k_gather<CT, AT, NZV, FL><<<nwg, 256, lds, s>>>(a, srt, w.perm, w.counts, ph, o);      \
    break;
    GSE_G_CASE(9)
    GSE_G_CASE(11)
    GSE_G_CASE(13)
    GSE_G_CASE(15)
    GSE_G_CASE(17)
#undef GSE_G_CASE
    default:
      k_gather<CT, AT, 0, FL><<<nwg, 256, lds, s>>>(a, srt, w.perm, w.counts, ph, o);
  }
  return hipGetLastError();
}

template <typename CT, typename AT>
static hipError_t gather_fl(int flags, const GatherArgs& a, const Binning& w, const void* phi,
                            const GatherOut& o, size_t lds, unsigned nwg, hipStream_t s) {
  // smallest instantiated superset of the requested outputs
  const int S = GF_SPHERE, C = GF_CUBE, P = GF_POT, V = GF_VIRIAL;
  if ((flags & ~S) == 0) return gather_nz<CT, AT, S>(a, w, phi, o, lds, nwg, s);
  if ((flags & ~(S | P)) == 0) return gather_nz<CT, AT, S | P>(a, w, phi, o, lds, nwg, s);
  if ((flags & ~(C | P)) == 0) return gather_nz<CT, AT, C | P>(a, w, phi, o, lds, nwg, s);
  if ((flags & ~(C | P | V)) == 0) return gather_nz<CT, AT, C | P | V>(a, w, phi, o, lds, nwg, s);
  return gather_nz<CT, AT, S | C | P | V>(a, w, phi, o, lds, nwg, s);
}

hipError_t launch_gather(bool ct_double, bool at_double, const Geom& g, const BinShape& b,
                         double sigma, double rcut, int flags, const Binning& w, int n,
                         const void* phi, const GatherOut& o, hipStream_t s) {
  (void)n;
  GatherArgs a;
  a.g = g;
  a.b = b;
  a.inv2s2 = 1.0 / (2.0 * sigma * sigma);
  a.rc2 = rcut * rcut;
  a.prefF = pow(2.0 * M_PI, -3.0 / 2.0) * g.hx * g.hy * g.hz * pow(sigma, -5.0);
  a.prefP = pow(2.0 * M_PI * sigma * sigma, -3.0 / 2.0) * g.hx * g.hy * g.hz;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const bool lane_ok = NX + NY + NZ <= 64 && (NX * NY + 63) / 64 <= kMaxK;
  if (b.direct || !lane_ok) {
    if (ct_double)
      k_gather_global<double, double><<<2048, 256, 0, s>>>(a, flags, (const V4<double>*)w.sorted,
                                                           w.perm, w.counts, (const double*)phi, o);
    else if (at_double)
      k_gather_global<float, double><<<2048, 256, 0, s>>>(a, flags, (const V4<float>*)w.sorted,
                                                          w.perm, w.counts, (const float*)phi, o);
    else
      k_gather_global<float, float><<<2048, 256, 0, s>>>(a, flags, (const V4<float>*)w.sorted,
                                                         w.perm, w.counts, (const float*)phi, o);
    return hipGetLastError();
  }
  const size_t lds = b.brick_bytes(ct_double ? 8 : 4);
  const unsigned nwg = (unsigned)(((b.nbins + 7) / 8) * 8);
  if (ct_double) return gather_fl<double, double>(flags, a, w, phi, o, lds, nwg, s);
  if (at_double) return gather_fl<float, double>(flags, a, w, phi, o, lds, nwg, s);
  return gather_fl<float, float>(flags, a, w, phi, o, lds, nwg, s);
}

}  // namespace gse
