// Grid -> particle gather: the reference force
// (CpuGaussCharge::interpolateLoop, CpuGaussCharge.cpp:105-172: sphere
// support r2 < rcut2 inside the (2n+1)^3 cube) and the build-defined
// cube-support outputs (exact gradient, per-particle potential, virial
// moments) in one pass over the cube.
//
// One workgroup per bin (the binning of the spread).  The bin's potential
// brick of (B+2n)^3 values and its particles are staged into LDS with
// coalesced loads; one wave per particle computes the 3(2n+1) offsets and
// Gaussian factors; then one wave per particle walks the cube, lanes on
// (tx,ty) columns walking tz with the z factors in registers, and the partial
// sums are reduced across the wave with a fixed butterfly.  kOverflowWG extra
// workgroups gather the overflow particles from global memory.
#include <math.h>

#include "gse_device.h"
#include "gse_internal.h"
#include "walk_plan.h"

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
  int cap;       // particles staged per chunk
  int cap_bin;   // slots per bin
  int nwg_bins;
  const int* order;  // bin of each workgroup position (NULL: the position itself)
  int* order_next;   // the schedule of this binning's counts (NULL: not computed)
  int order_wg0;     // first of the kOrderWG workgroups that compute it
};

__device__ __forceinline__ float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmac(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Writes one particle's outputs from its (wave-reduced) sums.
template <typename AT, int FL>
__device__ __forceinline__ void gather_store(const GatherArgs& a, const GatherOut& o, int oi,
                                             double qv, double ds, const AT (&f)[13]) {
  const double fq = qv * a.prefF * o.fscale * ds;
  const double vq = qv * a.prefF * o.vscale;
  if ((FL & GF_SPHERE) && o.fx) {
    ((AT*)o.fx)[oi] = (AT)(fq * (double)f[0]);
    ((AT*)o.fy)[oi] = (AT)(fq * (double)f[1]);
    ((AT*)o.fz)[oi] = (AT)(fq * (double)f[2]);
  }
  if ((FL & GF_CUBE) && o.cx) {
    ((AT*)o.cx)[oi] = (AT)(fq * (double)f[3]);
    ((AT*)o.cy)[oi] = (AT)(fq * (double)f[4]);
    ((AT*)o.cz)[oi] = (AT)(fq * (double)f[5]);
  }
  if ((FL & GF_POT) && o.pot) ((AT*)o.pot)[oi] = (AT)(a.prefP * o.pscale * ds * (double)f[6]);
  if ((FL & GF_VIRIAL) && o.vir6) {
    const size_t st = (size_t)o.vir_stride;
    for (int c = 0; c < 6; c++) o.vir6[c * st + oi] = vq * (double)f[7 + c];
  }
}

// One particle gathered straight from global memory, one wave: used for the
// overflow list and for supports too wide for the LDS path.
template <typename CT, typename AT, int FL>
__device__ void gather_one_global(const GatherArgs& a, const GatherOut& o, const V4<CT>& p,
                                  int oi, double ds, const CT* __restrict__ phi) {
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1, NZ = 2 * g.nz + 1;
  const CT inv2s2 = (CT)a.inv2s2;
  const CT rc2 = (CT)a.rc2;
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
  if (lane == 0) gather_store<AT, FL>(a, o, oi, (double)p.w, ds, f);
}

// LDS layout of k_gather: staged 4-vectors (32-byte aligned) at 0, then the
// factor pairs (w, d) of type CT, q (double), brick offsets and original
// indices (int), then the potential brick (CT).
struct GatherLds {
  size_t oX, oW, oQ, oB, oP, oBrick, total;
};
__host__ __device__ inline GatherLds gather_lds(int vol, int cap, int nw, int ct_bytes) {
  GatherLds l;
  size_t o = 0;
  l.oX = o;
  o += (size_t)cap * 4 * ct_bytes;
  l.oW = o;
  o += (size_t)cap * nw * 2 * ct_bytes;
  o = (o + 7) & ~(size_t)7;
  l.oQ = o;
  o += (size_t)cap * 8;
  l.oB = o;
  o += (size_t)cap * 4;
  l.oP = o;
  o += (size_t)cap * 4;
  o = (o + 15) & ~(size_t)15;
  l.oBrick = o;
  o += (size_t)vol * ct_bytes;
  l.total = o;
  return l;
}

// Row pitch of the staged potential brick: FP32 bricks on grids whose rows
// split into aligned 16-byte pieces (X and the bin edge multiples of 4) are
// staged with 16-byte loads and stores, so their rows are padded to a
// multiple of 4 values.
__host__ __device__ inline bool brick_v4(int ct_bytes, int X, int Bx) {
  return ct_bytes == 4 && X % 4 == 0 && Bx % 4 == 0;
}
__host__ __device__ inline int brick_pitch(int ct_bytes, int X, int Bx, int Lx) {
  return brick_v4(ct_bytes, X, Bx) ? (Lx + 3) & ~3 : Lx;
}

// FL: GatherFlags bit mask; NZC: 2nz+1 known at compile time (0 = runtime);
// PL: the brick's plane size Px*Ly known at compile time (0 = runtime), so
// the cube walk's LDS reads take their z step as an immediate offset
//
// One bin of one system: its cnt particles start at `sorted`/`perm` (slot 0
// of the bin), the system's potential grid at `phi`; perm holds the output
// index of each particle.  Ends with a barrier, so a workgroup may call it
// for several bins in a row.
template <typename CT, typename AT, int NZC, int FL, int BS, int PL = 0, int GS = 32>
__device__ __forceinline__ void gather_bin(const GatherArgs& a, int bin, int cnt,
                                           const V4<CT>* __restrict__ sorted,
                                           const int* __restrict__ perm,
                                           const int* __restrict__ loc,
                                           const CT* __restrict__ phi, const GatherOut& o,
                                           double ds, double* smem_raw) {
  const Geom& g = a.g;
  const BinShape& b = a.b;
  const int s0 = 0, s1 = cnt;
  const int bxi = bin % b.nbx, byi = (bin / b.nbx) % b.nby, bzi = bin / (b.nbx * b.nby);
  const int Lx = b.Lx, Ly = b.Ly, Lz = b.Lz;
  const int Px = brick_pitch(sizeof(CT), g.X, b.Bx, Lx), PxLy = Px * Ly;
  const int NX = 2 * g.nx + 1, NY = 2 * g.ny + 1;
  const int NZ = NZC > 0 ? NZC : 2 * g.nz + 1;
  const int NW = NX + NY + NZ;
  const int cap = a.cap;
  const GatherLds lay = gather_lds(PxLy * Lz, cap, NW, sizeof(CT));
  char* lds = (char*)smem_raw;
  V4<CT>* sX = (V4<CT>*)(lds + lay.oX);
  CT* sW = (CT*)(lds + lay.oW);  // [cap][NW][2] (w, d)
  double* sQ = (double*)(lds + lay.oQ);
  int* sB = (int*)(lds + lay.oB);
  int* sP = (int*)(lds + lay.oP);
  CT* brick = (CT*)(lds + lay.oBrick);

  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int nwv = BS / 64;
  // the first chunk's particles (cap <= 256 <= BS: at most one per thread)
  // are loaded first, so their latency overlaps the brick staging
  // (the index is clamped into the bin's slots, so the loads need no branch
  // and the values stay in registers; slots past cnt are never used)
  const int pslot = min((int)threadIdx.x, a.cap_bin - 1);
  const V4<CT> pre = sorted[pslot];
  const int pre_p = perm[pslot], pre_l = loc[pslot];
  {  // stage the potential brick: rows of Lx lanes, no per-element division.
     // Loads go out in batches of kStageBatch before their LDS stores, so a
     // thread has that many in flight instead of one.  Wrapped grid rows come
     // from two small tables (Ly, Lz <= 64: the brick may be wider than the
     // grid, so one conditional wrap would not do).
    __shared__ int s_gy[64], s_gz[64];
    if (threadIdx.x < (unsigned)Ly) {
      int v = byi * b.By + (int)threadIdx.x;
      while (v >= g.Y) v -= g.Y;
      s_gy[threadIdx.x] = v;
    } else if (threadIdx.x >= 64 && threadIdx.x < 64u + (unsigned)Lz) {
      int v = bzi * b.Bz + (int)threadIdx.x - 64;
      while (v >= g.Z) v -= g.Z;
      s_gz[threadIdx.x - 64] = v;
    }
    __syncthreads();
    constexpr int kStageBatch = 8;
    if (sizeof(CT) == 4 && brick_v4(4, g.X, b.Bx)) {
      // 16-byte pieces: row r = ly + Ly*lz, piece j covers columns 4j..4j+3
      const int nq = Px >> 2, xq = g.X >> 2, gq0 = (bxi * b.Bx) >> 2;
      const int total = Ly * Lz * nq;
      const float inv_nq = 1.0f / (float)nq, inv_ly = 1.0f / (float)Ly;
      for (int e0 = threadIdx.x; e0 < total; e0 += kStageBatch * BS) {
        float4 v[kStageBatch];
        int dst[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; u++) {
          const int e = e0 + u * BS;
          const bool ok = e < total;
          const int row = (int)(((float)e + 0.5f) * inv_nq);
          const int j = e - row * nq;
          int lz = (int)(((float)row + 0.5f) * inv_ly);
          const int ly = row - lz * Ly;
          lz = lz < Lz ? lz : Lz - 1;
          const int qx = (gq0 + j) % xq;
          dst[u] = ok ? Px * row + 4 * j : -1;
          v[u] = ok ? ((const float4*)phi)[qx + (size_t)xq * (s_gy[ly] + (size_t)g.Y * s_gz[lz])]
                    : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; u++)
          if (dst[u] >= 0) *(float4*)&brick[dst[u]] = v[u];
      }
    } else {
    const int rows_per = 64 / Lx;
    const int rl = lane / Lx;
    const int lx = lane - rl * Lx;
    const bool row_lane = rl < rows_per;
    int gx = bxi * b.Bx + lx;
    while (gx >= g.X) gx -= g.X;
    const int rows = Ly * Lz;
    const int rstep = nwv * rows_per;
    const int rstep_z = rstep / Ly, rstep_y = rstep - rstep_z * Ly;
    int r = wid * rows_per + rl;
    int ly = r % Ly, lz = r / Ly;
    for (; r < rows; r += kStageBatch * rstep) {
      CT v[kStageBatch];
      int dst[kStageBatch];
#pragma unroll
      for (int u = 0; u < kStageBatch; u++) {
        const bool ok = row_lane && r + u * rstep < rows;
        const int tz = lz < Lz ? lz : Lz - 1;  // rows past the end: any valid entry
        dst[u] = ok ? lx + Px * (ly + Ly * lz) : -1;
        v[u] = ok ? phi[gx + (size_t)g.X * (s_gy[ly] + (size_t)g.Y * s_gz[tz])] : (CT)0;
        ly += rstep_y;
        lz += rstep_z;
        if (ly >= Ly) {
          ly -= Ly;
          lz++;
        }
      }
#pragma unroll
      for (int u = 0; u < kStageBatch; u++)
        if (dst[u] >= 0) brick[dst[u]] = v[u];
    }
    }
  }

  const int NXY = NX * NY;
  const int KI = (NXY + 63) >> 6;
  const CT rc2 = (CT)a.rc2;
  const CT inv2s2 = (CT)a.inv2s2;
  double hsel, ihsel;
  int nsel, jj, axis;
  // the half-widths as opaque uniform values: a per-lane choice between
  // fields of the argument struct would otherwise become an indexed copy
  // of the struct in scratch memory
  const int gnx = __builtin_amdgcn_readfirstlane(g.nx);
  const int gny = __builtin_amdgcn_readfirstlane(g.ny);
  const int gnz = __builtin_amdgcn_readfirstlane(g.nz);
  // the spacings as wave-uniform values (the paired form selects per lane)
  const double hx_u = g.hx, hy_u = g.hy, hz_u = g.hz;
  const double ihx_u = g.ihx, ihy_u = g.ihy, ihz_u = g.ihz;
  if (lane < NX) {
    axis = 0; jj = lane; hsel = g.hx; ihsel = g.ihx; nsel = gnx;
  } else if (lane < NX + NY) {
    axis = 1; jj = lane - NX; hsel = g.hy; ihsel = g.ihy; nsel = gny;
  } else {
    axis = lane < NW ? 2 : -1; jj = lane - NX - NY; hsel = g.hz; ihsel = g.ihz; nsel = gnz;
  }
  int k_tx[kMaxK], k_ty[kMaxK], k_off[kMaxK];
  bool k_ok[kMaxK];
#pragma unroll
  for (int k = 0; k < kMaxK; k++) {
    const int idx = lane + 64 * k;
    k_ok[k] = (k < KI) && idx < NXY;
    k_ty[k] = k_ok[k] ? idx / NX : 0;
    k_tx[k] = k_ok[k] ? idx - k_ty[k] * NX : 0;
    k_off[k] = k_tx[k] + Px * k_ty[k];
  }

  constexpr bool SPH = (FL & GF_SPHERE) != 0;
  constexpr bool CUB = (FL & GF_CUBE) != 0;
  constexpr bool POT = (FL & GF_POT) != 0;
  constexpr bool VIR = (FL & GF_VIRIAL) != 0;
  constexpr bool NEED_B0 = CUB || POT || VIR;
  constexpr bool NEED_B1 = CUB || VIR;
  constexpr int NZA = NZC > 0 ? NZC : 1;
  // FP32 without the virial: the sphere test dx^2+dy^2 < rc^2-dz^2 with the
  // right side kept per z (one compare per point).  FP64 keeps the
  // reference's rounding, (dx^2+dy^2)+dz^2 < rc^2 (CpuGaussCharge.cpp:157).
  constexpr bool FAST_SPH = SPH && !VIR && sizeof(CT) == 4;
  // FP32 accumulation: the z walk as FMAs (see below)
  constexpr bool FMA_FORM = NZC > 0 && sizeof(CT) == sizeof(AT);
  const int PLv = PL > 0 ? PL : PxLy;
  // PL > 0 also means a cubic support of NZC points per axis.  The FMA form
  // then walks two charges per wave, one per 32-lane half: the per-charge
  // set-up and the cross-lane reduction are shared by two charges, and the
  // NZC^2 columns fill 32-lane slots instead of 64-lane ones.
  // FP64 pairs too (without the virial sums, whose extra z tables would not
  // fit beside two charges' registers), with the reference's exact test
  constexpr bool PAIR = FMA_FORM && PL > 0 && (sizeof(CT) == 4 || !VIR);
  constexpr bool EXACT_SPH = VIR || sizeof(CT) == 8;
  // GS lanes per charge (32: two charges per wave; 16: four)
  static_assert(GS == 32 || GS == 16, "charge groups of 32 or 16 lanes");
  constexpr int GPW = 64 / GS;
  constexpr int NXYP = NZA * NZA, KP = (NXYP + GS - 1) / GS;
  // the paired walk's column order and sphere classes (PL > 0: the support
  // is cubic; PL = 400 and 528 are 20 x 20 and 24 x 22 brick planes)
  using Plan = WalkPlanOf<NZA, GS, PL == 400 ? 20 : PL == 528 ? 24 : 0>;
  int p_tx[KP], p_ty[KP];
#pragma unroll
  for (int k = 0; k < KP; k++) {
    const int idx = (lane & (GS - 1)) + GS * k;
    const int c = PAIR && idx < NXYP ? Plan::v.col[idx] : idx < NXYP ? idx : 0;
    p_ty[k] = c / NZA;
    p_tx[k] = c - p_ty[k] * NZA;
  }

  for (int c0 = s0; c0 < s1; c0 += cap) {
    const int na = min(cap, s1 - c0);
    // phase 0: particles and original indices into LDS, one coalesced pass
    if ((int)threadIdx.x < na) {
      CT px = pre.x, py = pre.y, pz = pre.z, pw = pre.w;
      int l = pre_l, pp = pre_p;
      if (c0 != 0) {
        const V4<CT> q = sorted[c0 + threadIdx.x];
        px = q.x, py = q.y, pz = q.z, pw = q.w;
        l = loc[c0 + threadIdx.x];
        pp = perm[c0 + threadIdx.x];
      }
      const V4<CT> p{px, py, pz, pw};
      sX[threadIdx.x] = p;
      sP[threadIdx.x] = pp;
      // brick offset of the cube origin (stored by k_bin) and the charge
      sB[threadIdx.x] = (l & 255) + Px * (((l >> 8) & 255) + Ly * (l >> 16));
      sQ[threadIdx.x] = (double)p.w;
    }
    __syncthreads();
    // phase 1: one wave per particle; lane j holds the reference gather's
    // offset d = x - t*h (CpuGaussCharge.cpp:146-155, t = round(x/h) - n + j)
    // and factor exp(-d^2/2s^2)
    if constexpr (PAIR && sizeof(CT) == 8) {
      // FP64, cubic support: the chunk's na x 3 NZA factors one per lane
      // over the whole workgroup, entry e = ai NW + j as above (in FP32 the
      // per-lane axis selects cost more VALU than the idle lanes saved)
      constexpr int NW3 = 3 * NZA;
      for (int e = threadIdx.x; e < na * NW3; e += BS) {
        const int ai = e / NW3, j = e - ai * NW3;
        const int ax = j / NZA, jx = j - ax * NZA;
        const V4<CT> p = sX[ai];
        const double ps = ax == 0 ? (double)p.x : ax == 1 ? (double)p.y : (double)p.z;
        const double hs = ax == 0 ? hx_u : ax == 1 ? hy_u : hz_u;
        const double ihs = ax == 0 ? ihx_u : ax == 1 ? ihy_u : ihz_u;
        const int c = (int)round(ps * ihs);
        const CT d = (CT)sub_rn(ps, mul_rn((double)(c - gnz + jx), hs));
        sW[2 * e] = gexp(mul_rn(mul_rn(-d, d), inv2s2));
        sW[2 * e + 1] = d;
      }
    } else {
      for (int ai = wid; ai < na; ai += nwv) {
        const V4<CT> p = sX[ai];
        const double ps = axis == 0 ? (double)p.x : axis == 1 ? (double)p.y : (double)p.z;
        if (axis >= 0) {
          const int c = (int)round(ps * ihsel);
          const CT d = (CT)sub_rn(ps, mul_rn((double)(c - nsel + jj), hsel));
          sW[2 * ((size_t)ai * NW + lane)] = gexp(mul_rn(mul_rn(-d, d), inv2s2));
          sW[2 * ((size_t)ai * NW + lane) + 1] = d;
        }
      }
    }
    __syncthreads();
    if constexpr (PAIR) {
      // phase 2, GPW charges per wave: lane group h of wave `wid` takes
      // charge GPW*g2 + h; a group without a charge walks the first group's
      // charge with its sums discarded (no divergence inside the walk)
      const int h = lane / GS, l32 = lane & (GS - 1);
      for (int g2 = wid; GPW * g2 < na; g2 += nwv) {
        const int ai = GPW * g2 + h;
        const bool live = ai < na;
        const int aic = live ? ai : GPW * g2;
        const CT* w = sW + (size_t)aic * NW * 2;
        const int base = sB[aic];
        // dz2: rc^2 - dz^2 (no virial: one compare per point), else dz^2
        // (the virial weight, and the exact sphere test); w2r = wz dz^2
        CT wzr[NZA], dz2[NZA], vzr[NZA], w2r[VIR ? NZA : 1];
#pragma unroll
        for (int t = 0; t < NZA; t++) {
          wzr[t] = w[2 * (2 * NZA + t)];
          const CT dz = w[2 * (2 * NZA + t) + 1];
          dz2[t] = EXACT_SPH ? mul_rn(dz, dz) : rc2 - mul_rn(dz, dz);
          vzr[t] = wzr[t] * dz;
          if constexpr (VIR) w2r[t] = wzr[t] * dz2[t];
        }
        AT f[13];
#pragma unroll
        for (int c = 0; c < 13; c++) f[c] = 0;
#pragma unroll
        for (int k = 0; k < KP; k++) {
          if (l32 + GS * k >= NXYP) continue;
          const CT wx = w[2 * p_tx[k]];
          const CT dx = w[2 * p_tx[k] + 1];
          const CT wy = w[2 * (NZA + p_ty[k])];
          const CT dy = w[2 * (NZA + p_ty[k]) + 1];
          const CT dxy2 = add_rn(mul_rn(dx, dx), mul_rn(dy, dy));
          const int a0 = base + p_tx[k] + Px * p_ty[k];
          AT B0 = 0, B1 = 0, B2 = 0, A0 = 0, A1 = 0;
#pragma unroll
          for (int t = 0; t < NZA; t++) {
            // 1: inside the sphere for every charge, 2: outside, 0: tested
            const int cl = Plan::v.cls[k][t];
            const CT ph = brick[a0 + t * PL];
            if (NEED_B0) B0 = fmac((AT)wzr[t], (AT)ph, B0);
            if (NEED_B1) B1 = fmac((AT)vzr[t], (AT)ph, B1);
            if constexpr (VIR) B2 = fmac((AT)w2r[t], (AT)ph, B2);
            if (SPH && cl == 1) {
              A0 = fmac((AT)wzr[t], (AT)ph, A0);
              A1 = fmac((AT)vzr[t], (AT)ph, A1);
            } else if (SPH && cl == 0) {
              // (dx^2 + dy^2) + dz^2 < rc^2 as CpuGaussCharge.cpp:157 rounds
              // it (FP64, virial), else dx^2 + dy^2 < rc^2 - dz^2
              const bool in = EXACT_SPH ? add_rn(dxy2, dz2[t]) < rc2 : dxy2 < dz2[t];
              const AT ms = in ? (AT)ph : (AT)0;
              A0 = fmac((AT)wzr[t], ms, A0);
              A1 = fmac((AT)vzr[t], ms, A1);
            }
          }
          const AT wxy = wx * wy;
          if (SPH) {
            const AT t0 = wxy * A0;
            f[0] = fmac(t0, dx, f[0]);
            f[1] = fmac(t0, dy, f[1]);
            f[2] = fmac(wxy, A1, f[2]);
          }
          if (NEED_B0) {
            const AT t0 = wxy * B0;
            if (CUB) {
              f[3] = fmac(t0, dx, f[3]);
              f[4] = fmac(t0, dy, f[4]);
              f[5] = fmac(wxy, B1, f[5]);
            }
            if (POT) f[6] += t0;
            if (VIR) {
              f[7] = fmac(t0 * dx, dx, f[7]);
              f[8] = fmac(t0 * dx, dy, f[8]);
              f[9] = fmac(wxy * dx, B1, f[9]);
              f[10] = fmac(t0 * dy, dy, f[10]);
              f[11] = fmac(wxy * dy, B1, f[11]);
              f[12] = fmac(wxy, B2, f[12]);
            }
          }
        }
        if constexpr (GS == 32 && sizeof(AT) == 4) {
          // FP32: the sums in fixed groups of four (sphere force and phi_i;
          // cube force; virial moments), each group over its 32-lane half in
          // one transposed butterfly (group4_total), so every output has
          // the same summation order whichever other outputs are requested;
          // lane k < 4 of the group stores entry k of each group
          const AT z0 = (AT)0;
          const AT sA = group4_total(live && SPH ? f[0] : z0, live && SPH ? f[1] : z0,
                                     live && SPH ? f[2] : z0, live && POT ? f[6] : z0, lane);
          AT sB = z0, sC = z0, sD = z0;
          if constexpr (CUB) sB = group4_total(live ? f[3] : z0, live ? f[4] : z0,
                                               live ? f[5] : z0, z0, lane);
          if constexpr (VIR) {
            sC = group4_total(live ? f[7] : z0, live ? f[8] : z0, live ? f[9] : z0,
                              live ? f[10] : z0, lane);
            sD = group4_total(live ? f[11] : z0, live ? f[12] : z0, z0, z0, lane);
          }
          if (live && l32 < 4) {
            const int oi = sP[ai];
            const double qv = sQ[ai];
            const double fq = qv * a.prefF * o.fscale * ds;
            if (SPH && l32 < 3 && o.fx) {
              AT* dst = (AT*)(l32 == 0 ? o.fx : l32 == 1 ? o.fy : o.fz);
              dst[oi] = (AT)(fq * (double)sA);
            }
            if (POT && l32 == 3 && o.pot)
              ((AT*)o.pot)[oi] = (AT)(a.prefP * o.pscale * ds * (double)sA);
            if (CUB && l32 < 3 && o.cx) {
              AT* dst = (AT*)(l32 == 0 ? o.cx : l32 == 1 ? o.cy : o.cz);
              dst[oi] = (AT)(fq * (double)sB);
            }
            if (VIR && o.vir6) {
              const double vq = qv * a.prefF * o.vscale;
              const size_t st = (size_t)o.vir_stride;
              o.vir6[l32 * st + oi] = vq * (double)sC;
              if (l32 < 2) o.vir6[(4 + l32) * st + oi] = vq * (double)sD;
            }
          }
        } else {
#pragma unroll
          for (int c = 0; c < 13; c++) {
            const bool used = (c < 3 && SPH) || (c >= 3 && c < 6 && CUB) || (c == 6 && POT) ||
                              (c >= 7 && VIR);
            if (used)
              f[c] = GS == 32 ? half_total(live ? f[c] : (AT)0) : row_total(live ? f[c] : (AT)0);
          }
          if (live && l32 == 0) gather_store<AT, FL>(a, o, sP[ai], sQ[ai], ds, f);
        }
      }
      __syncthreads();
      continue;
    }
    // phase 2: one wave per particle
    for (int ai = wid; ai < na; ai += nwv) {
      const CT* w = sW + (size_t)ai * NW * 2;
      const int base = sB[ai];
      CT wzr[NZA], dzr[NZA], dz2[NZA], vzr[NZA];
      if (NZC > 0) {
#pragma unroll
        for (int t = 0; t < NZA; t++) {
          wzr[t] = w[2 * (NX + NY + t)];
          dzr[t] = w[2 * (NX + NY + t) + 1];
          dz2[t] = FAST_SPH ? rc2 - mul_rn(dzr[t], dzr[t]) : mul_rn(dzr[t], dzr[t]);
          vzr[t] = wzr[t] * dzr[t];
        }
      }
      AT f[13];
#pragma unroll
      for (int c = 0; c < 13; c++) f[c] = 0;
#pragma unroll
      for (int k = 0; k < kMaxK; k++) {
        if (k >= KI) break;
        if (!k_ok[k]) continue;
        const CT wx = w[2 * k_tx[k]];
        const CT dx = w[2 * k_tx[k] + 1];
        const CT wy = w[2 * (NX + k_ty[k])];
        const CT dy = w[2 * (NX + k_ty[k]) + 1];
        // r2 = (dx*dx + dy*dy) + dz*dz, each product rounded (:157)
        const CT dxy2 = add_rn(mul_rn(dx, dx), mul_rn(dy, dy));
        const int a0 = base + k_off[k];
        AT B0 = 0, B1 = 0, B2 = 0, A0 = 0, A1 = 0;
        if (FMA_FORM && !VIR) {
          // each potential value enters every sum through one FMA against a
          // per-particle z factor (wz, wz*dz), the sphere mask selects it;
          // FP64 keeps the reference's exact sphere test
#pragma unroll
          for (int t = 0; t < NZA; t++) {
            const CT ph = brick[a0 + t * PLv];
            if (NEED_B0) B0 = fmac((AT)wzr[t], (AT)ph, B0);
            if (NEED_B1) B1 = fmac((AT)vzr[t], (AT)ph, B1);
            if (SPH) {
              const bool in = FAST_SPH ? dxy2 < dz2[t] : add_rn(dxy2, dz2[t]) < rc2;
              const AT ms = in ? (AT)ph : (AT)0;
              A0 = fmac((AT)wzr[t], ms, A0);
              A1 = fmac((AT)vzr[t], ms, A1);
            }
          }
        } else if (NZC > 0) {
#pragma unroll
          for (int t = 0; t < NZA; t++) {
            const CT ph = brick[a0 + t * PLv];
            const AT wph = (AT)(wzr[t] * ph);
            const AT dzt = (AT)dzr[t];
            if (NEED_B0) B0 += wph;
            if (NEED_B1) B1 = fmac(wph, dzt, B1);
            if (VIR) B2 = fmac(wph, (AT)dz2[t], B2);
            if (SPH) {
              const bool in = FAST_SPH ? dxy2 < dz2[t] : add_rn(dxy2, dz2[t]) < rc2;
              const AT ws = in ? wph : (AT)0;
              A0 += ws;
              A1 = fmac(ws, dzt, A1);
            }
          }
        } else {
          for (int t = 0; t < NZ; t++) {
            const CT wzt = w[2 * (NX + NY + t)];
            const CT dzc = w[2 * (NX + NY + t) + 1];
            const CT dzc2 = mul_rn(dzc, dzc);
            const CT ph = brick[a0 + t * PxLy];
            const AT wph = (AT)(wzt * ph);
            const AT dzt = (AT)dzc;
            if (NEED_B0) B0 += wph;
            if (NEED_B1) B1 = fmac(wph, dzt, B1);
            if (VIR) B2 = fmac(wph, (AT)dzc2, B2);
            if (SPH) {
              const AT ws = add_rn(dxy2, dzc2) < rc2 ? wph : (AT)0;
              A0 += ws;
              A1 = fmac(ws, dzt, A1);
            }
          }
        }
        const AT wxy = (AT)wx * (AT)wy;
        const AT ax = (AT)dx, ay = (AT)dy;
        if (SPH) {
          const AT t0 = wxy * A0;
          f[0] = fmac(t0, ax, f[0]);
          f[1] = fmac(t0, ay, f[1]);
          f[2] = fmac(wxy, A1, f[2]);
        }
        if (NEED_B0) {
          const AT t0 = wxy * B0;
          if (CUB) {
            f[3] = fmac(t0, ax, f[3]);
            f[4] = fmac(t0, ay, f[4]);
            f[5] = fmac(wxy, B1, f[5]);
          }
          if (POT) f[6] += t0;
          if (VIR) {
            f[7] = fmac(t0 * ax, ax, f[7]);
            f[8] = fmac(t0 * ax, ay, f[8]);
            f[9] = fmac(wxy * ax, B1, f[9]);
            f[10] = fmac(t0 * ay, ay, f[10]);
            f[11] = fmac(wxy * ay, B1, f[11]);
            f[12] = fmac(wxy, B2, f[12]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 13; c++) {
        const bool used = (c < 3 && SPH) || (c >= 3 && c < 6 && CUB) || (c == 6 && POT) ||
                          (c >= 7 && VIR);
        if (used) f[c] = wave_total(f[c]);
      }
      if (lane == 0) gather_store<AT, FL>(a, o, sP[ai], sQ[ai], ds, f);
    }
    __syncthreads();
  }
}


template <typename CT, typename AT, int NZC, int FL, int BS, int PL = 0, int GS = 32>
// FP32 with supports up to 13: at most 128 VGPRs, so two 512-thread
// workgroups share a CU
__global__ void __launch_bounds__(
    BS, (BS == 512 && sizeof(CT) == 4 && sizeof(AT) == 4 && NZC <= 13) ? 4 : 1)
    k_gather(GatherArgs a, const V4<CT>* __restrict__ xyzq, const V4<CT>* __restrict__ sorted,
             const int* __restrict__ perm, const int* __restrict__ counts,
             int* __restrict__ counts_other, const BinStats* __restrict__ stats,
             BinStats* __restrict__ stats_other, const int* __restrict__ ovf,
             const CT* __restrict__ phi, GatherOut o, const int* __restrict__ loc) {
  extern __shared__ __align__(32) double smem_raw[];
  const double ds = o.dscale ? *o.dscale : 1.0;
  if (a.order_next && (int)blockIdx.x >= a.order_wg0) {  // spare workgroups: the schedule
    bin_order_eighth<BS>(counts, a.b.nbins, (int)blockIdx.x - a.order_wg0, a.order_next,
                         (unsigned*)smem_raw);
    return;
  }
  if ((int)blockIdx.x == a.nwg_bins + kOverflowWG) {  // spare workgroup: k-pass finalize
    kpass_finalize_block(o.fin_partials, o.fin_nblocks, o.fin_ncomp, o.fin_pref, o.fin_energy,
                         o.fin_virial);
    return;
  }
  if ((int)blockIdx.x >= a.nwg_bins) {  // overflow particles from global memory
    if (blockIdx.x == (unsigned)a.nwg_bins && threadIdx.x == 0) {
      stats_other->qmax_bits = 0ull;
      stats_other->invalid = 0;
      stats_other->overflow = 0;
      stats_other->spill = 0;
    }
    const int novf = stats->overflow;
    const int gw = (blockIdx.x - a.nwg_bins) * (BS / 64) + (threadIdx.x >> 6);
    for (int k = __builtin_amdgcn_readfirstlane(gw); k < novf; k += kOverflowWG * (BS / 64)) {
      const int oi = ovf[k];
      gather_one_global<CT, AT, FL>(a, o, xyzq[oi], oi, ds, phi);
    }
    return;
  }
  const int pos = xcd_remap(blockIdx.x, a.nwg_bins);
  if (pos >= a.b.nbins) return;
  const int bin = a.order ? a.order[pos] : pos;
  const int cnt = min(counts[bin], a.cap_bin);
  if (threadIdx.x == 0) counts_other[bin] = 0;
  if (cnt == 0) return;
  const size_t s0 = (size_t)bin * a.cap_bin;
  gather_bin<CT, AT, NZC, FL, BS, PL, GS>(a, bin, cnt, sorted + s0, perm + s0, loc + s0, phi, o, ds,
                                  smem_raw);
}

// Supports too wide for the LDS path: every binned particle (one bin of
// capacity n) from global memory, one wave per particle.
template <typename CT, typename AT, int FL>
__global__ void __launch_bounds__(256)
    k_gather_direct(GatherArgs a, const V4<CT>* __restrict__ sorted, const int* __restrict__ perm,
                    const int* __restrict__ counts, int* __restrict__ counts_other,
                    BinStats* __restrict__ stats_other, const CT* __restrict__ phi, GatherOut o) {
  if (blockIdx.x == gridDim.x - 1) {  // spare workgroup: k-pass finalize
    if (o.fin_partials)
      kpass_finalize_block(o.fin_partials, o.fin_nblocks, o.fin_ncomp, o.fin_pref, o.fin_energy,
                           o.fin_virial);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts_other[0] = 0;
    stats_other->qmax_bits = 0ull;
    stats_other->invalid = 0;
    stats_other->overflow = 0;
    stats_other->spill = 0;
  }
  const double ds = o.dscale ? *o.dscale : 1.0;
  const int n = min(counts[0], a.cap_bin);
  // every block but the spare one shares the particles
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = ((gridDim.x - 1) * blockDim.x) >> 6;
  for (int i = gw; i < n; i += nwave) gather_one_global<CT, AT, FL>(a, o, sorted[i], perm[i], ds, phi);
}

// ---------------------------------------------------------------------------
// Batched independent systems (see the batched spread): one launch gathers
// every system; workgroups walk the (system, bin) item table.
__device__ __forceinline__ int batch_system_g(const int* __restrict__ pofs, int nsys, int i) {
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

__device__ __forceinline__ GatherArgs batch_gather_args(const BatchSys& S) {
  GatherArgs a;
  a.g = S.g;
  a.b = S.b;
  a.inv2s2 = S.inv2s2;
  a.rc2 = S.rc2;
  a.prefF = S.prefF;
  a.prefP = S.prefP;
  a.cap = S.gather_cap;
  a.cap_bin = S.cap_bin;
  a.nwg_bins = 0;
  a.order = nullptr;
  a.order_next = nullptr;
  a.order_wg0 = 0;
  return a;
}

template <typename CT, typename AT, int NZC, int FL, int BS, int PL = 0>
__global__ void __launch_bounds__(BS)
    k_batch_gather(BatchDev d, const BatchSys* __restrict__ sysp, const int2* __restrict__ items,
                   int nitems, const CT* __restrict__ phi, GatherOut o) {
  extern __shared__ __align__(32) double smem_raw[];
  const int nx = (int)(gridDim.x >> 3);
  const int x = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3);
  const int chunk = (nitems + 7) >> 3;
  const int end = min(nitems, x * chunk + chunk);
  for (int it = x * chunk + j; it < end; it += nx) {
    const int2 item = items[it];
    const BatchSys& S = sysp[item.x];
    const int bin = item.y;
    const int cnt = min(d.counts[S.bofs + bin], S.cap_bin);
    if (cnt == 0) continue;
    const long long s0 = S.slot0 + (long long)bin * S.cap_bin;
    gather_bin<CT, AT, NZC, FL, BS, PL>(batch_gather_args(S), bin, cnt,
                                    (const V4<CT>*)d.sorted + s0, d.perm + s0, d.loc + s0,
                                    phi + S.gofs, o, 1.0, smem_raw);
  }
}

template <typename CT, typename AT, int FL>
__global__ void __launch_bounds__(256)
    k_batch_gather_ovf(BatchDev d, const V4<CT>* __restrict__ xyzq, const CT* __restrict__ phi,
                       GatherOut o) {
  const int novf = *d.novf;
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int k = gw; k < novf; k += nw) {
    const int i = d.ovf[k];
    const BatchSys& S = d.sys[batch_system_g(d.pofs, d.nsys, i)];
    gather_one_global<CT, AT, FL>(batch_gather_args(S), o, xyzq[i], i, 1.0, phi + S.gofs);
  }
}

template <typename CT, typename AT, int FL>
static hipError_t batch_gather_t(int nz, int bs, size_t lds, int grid, const BatchDev& d,
                                 const int2* items, int nitems, const void* phi,
                                 const GatherOut& o, bool pair, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const unsigned g = (unsigned)grid;
  const CT* ph = (const CT*)phi;
  if constexpr (sizeof(CT) == 4 && sizeof(AT) == 4 && !(FL & GF_VIRIAL)) {
    // every system of the group: cubic support, 8^3 bins, 4-aligned brick rows
    if (pair && nz == 13 && bs == 512) {  // 20 x 20 brick planes
      k_batch_gather<CT, AT, 13, FL, 512, 400><<<g, 512, lds, s>>>(d, d.sys, items, nitems, ph, o);
      return hipGetLastError();
    }
    if (pair && nz == 15 && bs == 512) {  // 24 x 22 brick planes
      k_batch_gather<CT, AT, 15, FL, 512, 528><<<g, 512, lds, s>>>(d, d.sys, items, nitems, ph, o);
      return hipGetLastError();
    }
  }
#define GSE_BG_LAUNCH(NZV, BSV) \
  k_batch_gather<CT, AT, NZV, FL, BSV><<<g, BSV, lds, s>>>(d, d.sys, items, nitems, ph, o)
#define GSE_BG_BS(NZV)       \
  if (bs == 512)             \
    GSE_BG_LAUNCH(NZV, 512); \
  else                       \
    GSE_BG_LAUNCH(NZV, 256);
  switch (nz) {
    case 13: GSE_BG_BS(13) break;
    case 15: GSE_BG_BS(15) break;
    case 17: GSE_BG_BS(17) break;
    default: GSE_BG_BS(0) break;
  }
#undef GSE_BG_BS
#undef GSE_BG_LAUNCH
  return hipGetLastError();
}

template <typename CT, typename AT>
static hipError_t batch_gather_at(int nz, int bs, size_t lds, int grid, const BatchDev& d,
                                  const int2* items, int nitems, const void* phi,
                                  const GatherOut& o, bool pair, hipStream_t s) {
  if (o.pot)
    return batch_gather_t<CT, AT, GF_SPHERE | GF_POT>(nz, bs, lds, grid, d, items, nitems, phi,
                                                      o, pair, s);
  return batch_gather_t<CT, AT, GF_SPHERE>(nz, bs, lds, grid, d, items, nitems, phi, o, pair, s);
}

hipError_t launch_batch_gather(bool ct_dbl, bool at_dbl, int nz, int block, size_t lds, int grid,
                               const BatchDev& d, const int2* items, int nitems, const void* phi,
                               const GatherOut& o, bool pair, hipStream_t s) {
  if (lds > 150 * 1024) return hipErrorInvalidConfiguration;
  if (ct_dbl)
    return batch_gather_at<double, double>(nz, block, lds, grid, d, items, nitems, phi, o, false,
                                           s);
  if (at_dbl)
    return batch_gather_at<float, double>(nz, block, lds, grid, d, items, nitems, phi, o, false,
                                          s);
  return batch_gather_at<float, float>(nz, block, lds, grid, d, items, nitems, phi, o, pair, s);
}

bool gather_pair_ok(bool ct_dbl, bool at_dbl, const Geom& g, const BinShape& b) {
  if (ct_dbl || at_dbl || g.nx != g.nz || g.ny != g.nz) return false;
  const int pl = brick_pitch(4, g.X, b.Bx, b.Lx) * b.Ly;
  return (g.nz == 6 && pl == 400) || (g.nz == 7 && pl == 528);
}

template <typename CT, typename AT>
static hipError_t batch_gather_ovf_t(const BatchDev& d, const void* xyzq, const void* phi,
                                     const GatherOut& o, hipStream_t s) {
  if (o.pot)
    k_batch_gather_ovf<CT, AT, GF_SPHERE | GF_POT><<<64, 256, 0, s>>>(
        d, (const V4<CT>*)xyzq, (const CT*)phi, o);
  else
    k_batch_gather_ovf<CT, AT, GF_SPHERE><<<64, 256, 0, s>>>(d, (const V4<CT>*)xyzq,
                                                            (const CT*)phi, o);
  return hipGetLastError();
}

hipError_t launch_batch_gather_overflow(bool ct_dbl, bool at_dbl, const BatchDev& d,
                                        const void* xyzq, const void* phi, const GatherOut& o,
                                        hipStream_t s) {
  if (ct_dbl) return batch_gather_ovf_t<double, double>(d, xyzq, phi, o, s);
  if (at_dbl) return batch_gather_ovf_t<float, double>(d, xyzq, phi, o, s);
  return batch_gather_ovf_t<float, float>(d, xyzq, phi, o, s);
}

size_t gather_lds_bytes(bool dbl, const Geom& g, const BinShape& b, int* cap_out) {
  const int NW = 2 * (g.nx + g.ny + g.nz) + 3;
  const int ct = dbl ? 8 : 4;
  const int vol = brick_pitch(ct, g.X, b.Bx, b.Lx) * b.Ly * b.Lz;
  const size_t lds_total = (size_t)vol * ct + (size_t)gather_stage_bytes(dbl);
  int cap = 256;
  while (cap > 1 && gather_lds(vol, cap, NW, ct).total > lds_total) cap--;
  if (cap_out) *cap_out = cap;
  return gather_lds(vol, cap, NW, ct).total;
}

// ---------------------------------------------------------------------------
template <typename CT, typename AT, int FL>
static hipError_t gather_nz(GatherArgs a, const Binning& w, const void* phi, const GatherOut& o,
                            hipStream_t s) {
  const V4<CT>* xq = (const V4<CT>*)w.xyzq;
  const V4<CT>* srt = (const V4<CT>*)w.sorted;
  const CT* ph = (const CT*)phi;
  if (a.b.direct) {
    a.nwg_bins = 1;
    k_gather_direct<CT, AT, FL><<<2049, 256, 0, s>>>(a, srt, w.perm, w.counts, w.counts_other,
                                                     w.stats_other, ph, o);
    return hipGetLastError();
  }
  const int NW = 2 * (a.g.nx + a.g.ny + a.g.nz) + 3;
  const int vol = brick_pitch(sizeof(CT), a.g.X, a.b.Bx, a.b.Lx) * a.b.Ly * a.b.Lz;
  const Tuning& tu = tuning();
  const size_t lds_total = (size_t)vol * sizeof(CT) + (size_t)gather_stage_bytes(sizeof(CT) == 8);
  int cap = 256;
  while (cap > 1 && gather_lds(vol, cap, NW, sizeof(CT)).total > lds_total) cap--;
  a.cap = cap;
  const size_t lds = gather_lds(vol, cap, NW, sizeof(CT)).total;
  if (lds > 150 * 1024) return hipErrorInvalidConfiguration;
  a.nwg_bins = ((a.b.nbins + 7) / 8) * 8;
  a.order = w.order;
  a.order_next = w.order_next;
  a.order_wg0 = a.nwg_bins + kOverflowWG + (o.fin_partials ? 1 : 0);
  const unsigned nwg = (unsigned)(a.order_wg0 + (a.order_next ? kOrderWG : 0));
  const int bs = tu.gather_block ? tu.gather_block : (sizeof(CT) == 4 ? 512 : 256);
#define GSE_G_ARGS a, xq, srt, w.perm, w.counts, w.counts_other, w.stats, w.stats_other, w.ovf, ph, o, \
                   w.loc
#define GSE_G_BS(NZV)                                                      \
  if (bs == 512)                                                           \
    k_gather<CT, AT, NZV, FL, 512><<<nwg, 512, lds, s>>>(GSE_G_ARGS);     \
  else                                                                     \
    k_gather<CT, AT, NZV, FL, 256><<<nwg, 256, lds, s>>>(GSE_G_ARGS);
  const int pxly = brick_pitch(sizeof(CT), a.g.X, a.b.Bx, a.b.Lx) * a.b.Ly;
  switch (2 * a.g.nz + 1) {
    case 11: GSE_G_BS(11) break;
    case 13:
      // 8^3 bins with n = 6 on every axis: 20 x 20 planes of the brick (FP32;
      // not the full output set, whose sums do not fit the register budget)
      if constexpr (sizeof(CT) == 4 && FL != (GF_SPHERE | GF_CUBE | GF_POT | GF_VIRIAL)) {
        if (bs == 512 && pxly == 400 && a.g.nx == a.g.nz && a.g.ny == a.g.nz) {
          if (tu.gather_group == 16)
            k_gather<CT, AT, 13, FL, 512, 400, 16><<<nwg, 512, lds, s>>>(GSE_G_ARGS);
          else
            k_gather<CT, AT, 13, FL, 512, 400><<<nwg, 512, lds, s>>>(GSE_G_ARGS);
          break;
        }
      }
      if constexpr (sizeof(CT) == 8 && sizeof(AT) == 8) {  // FP64: the z step as an immediate
        if (bs == 256 && pxly == 400 && a.g.nx == a.g.nz && a.g.ny == a.g.nz) {
          k_gather<CT, AT, 13, FL, 256, 400><<<nwg, 256, lds, s>>>(GSE_G_ARGS);
          break;
        }
      }
      GSE_G_BS(13)
      break;
    case 15: GSE_G_BS(15) break;
    case 17: GSE_G_BS(17) break;
    default: GSE_G_BS(0) break;
  }
#undef GSE_G_BS
#undef GSE_G_ARGS
  return hipGetLastError();
}

template <typename CT, typename AT>
static hipError_t gather_fl(int flags, const GatherArgs& a, const Binning& w, const void* phi,
                            const GatherOut& o, hipStream_t s) {
  // smallest instantiated superset of the requested outputs
  const int S = GF_SPHERE, C = GF_CUBE, P = GF_POT, V = GF_VIRIAL;
  if ((flags & ~S) == 0) return gather_nz<CT, AT, S>(a, w, phi, o, s);
  if ((flags & ~(S | P)) == 0) return gather_nz<CT, AT, S | P>(a, w, phi, o, s);
  if ((flags & ~(C | P)) == 0) return gather_nz<CT, AT, C | P>(a, w, phi, o, s);
  if ((flags & ~(C | P | V)) == 0) return gather_nz<CT, AT, C | P | V>(a, w, phi, o, s);
  return gather_nz<CT, AT, S | C | P | V>(a, w, phi, o, s);
}

hipError_t launch_gather(bool ct_double, bool at_double, const Geom& g, const BinShape& b,
                         double sigma, double rcut, int flags, const Binning& w, const void* phi,
                         const GatherOut& o, hipStream_t s) {
  GatherArgs a;
  a.g = g;
  a.b = b;
  a.inv2s2 = 1.0 / (2.0 * sigma * sigma);
  a.rc2 = rcut * rcut;
  a.prefF = pow(2.0 * M_PI, -3.0 / 2.0) * g.hx * g.hy * g.hz * pow(sigma, -5.0);
  a.prefP = pow(2.0 * M_PI * sigma * sigma, -3.0 / 2.0) * g.hx * g.hy * g.hz;
  a.cap = 0;
  a.cap_bin = w.cap_bin;
  a.nwg_bins = 0;
  a.order = nullptr;
  a.order_next = nullptr;
  a.order_wg0 = 0;
  if (ct_double) return gather_fl<double, double>(flags, a, w, phi, o, s);
  if (at_double) return gather_fl<float, double>(flags, a, w, phi, o, s);
  return gather_fl<float, float>(flags, a, w, phi, o, s);
}

}  // namespace gse
