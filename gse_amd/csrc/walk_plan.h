// Compile-time plan of the gather's paired cube walk (plain C++17
// constexpr, shared by the HIP kernels and a host-side check).
#pragma once

namespace gse {

// Static plan of the paired cube walk: N = 2n+1 points per axis, GS lanes
// per charge, slot k of lane l walking column col[l + GS k] (tx + N ty).
//
// The cube is centred on the grid point nearest the charge, so the charge
// sits at most h/2 from that point along each axis, and n = ceil(rcut/h) on
// every axis puts rcut/h in (n-1, n].  A cube point at offsets (i, j, k)
// cells from the centre is therefore inside the sphere for every charge if
// sum (|i|+1/2)^2 < (n-1)^2, and outside it for every charge if
// sum max(|i|-1/2, 0)^2 > n^2 (in units of h^2 of each axis; the sums are
// multiples of 1/4, so either bound holds with a margin of h^2/4 or more,
// far beyond rounding).  cls[k][t] records, for slot k and z step t, whether
// that holds for all the slot's columns (1 inside, 2 outside, 0 tested per
// point): the walk then skips the sphere test, or the sphere sums, there.
//
// The columns are taken in order of distance from the centre, so a slot's
// columns share their classes over long runs of z steps, with a greedy
// placement that keeps the brick offsets tx + PXC ty of one slot distinct
// modulo 64 words where the remaining columns allow it, as the row order
// does (PXC: the brick row pitch, 0 for the row order itself).
template <int N, int GS>
struct WalkPlan {
  static constexpr int NC = N * N, KS = (NC + GS - 1) / GS;
  int col[NC];
  int cls[KS][N];
};

constexpr int walk_abs(int v) { return v < 0 ? -v : v; }
constexpr int walk_sq(int v) { return v * v; }

template <int N, int GS, int PXC>
constexpr WalkPlan<N, GS> make_walk_plan() {
  WalkPlan<N, GS> p{};
  constexpr int n = N / 2, NC = N * N;
  int key[NC] = {}, ord[NC] = {};
  for (int c = 0; c < NC; c++) {
    const int i = c % N - n, j = c / N - n;
    // distance from the centre, then |i|, then (i, j)
    key[c] = (((i * i + j * j) * (n + 1) + walk_abs(i)) * N + i + n) * N + j + n;
    ord[c] = c;
  }
  for (int a = 1; a < NC; a++)
    for (int b = a; b > 0 && key[ord[b]] < key[ord[b - 1]]; b--) {
      const int x = ord[b];
      ord[b] = ord[b - 1];
      ord[b - 1] = x;
    }
  if (PXC > 0) {
    bool used[NC] = {};
    int pos = 0;
    while (pos < NC) {
      bool bank[64] = {};
      const int want = NC - pos < GS ? NC - pos : GS;
      int in_slot = 0;
      for (int q = 0; q < NC && in_slot < want; q++) {
        const int c = ord[q];
        const int r = (c % N + PXC * (c / N)) % 64;
        if (used[c] || bank[r]) continue;
        bank[r] = used[c] = true;
        p.col[pos++] = c;
        in_slot++;
      }
      // no distinct word left: the slot takes the next columns as they come
      for (int q = 0; q < NC && in_slot < want; q++) {
        const int c = ord[q];
        if (used[c]) continue;
        used[c] = true;
        p.col[pos++] = c;
        in_slot++;
      }
    }
  } else {
    for (int c = 0; c < NC; c++) p.col[c] = c;
  }
  for (int k = 0; k < p.KS; k++)
    for (int t = 0; t < N; t++) {
      bool all_in = true, all_out = true;
      for (int l = 0; l < GS && GS * k + l < NC; l++) {
        const int c = p.col[GS * k + l];
        const int o[3] = {c % N - n, c / N - n, t - n};
        int mx = 0, mn = 0;  // in units of h^2/4
        for (int a = 0; a < 3; a++) {
          mx += walk_sq(2 * walk_abs(o[a]) + 1);
          mn += o[a] != 0 ? walk_sq(2 * walk_abs(o[a]) - 1) : 0;
        }
        if (!(mx < 4 * (n - 1) * (n - 1))) all_in = false;
        if (!(mn > 4 * n * n)) all_out = false;
      }
      p.cls[k][t] = all_in ? 1 : all_out ? 2 : 0;
    }
  return p;
}

template <int N, int GS, int PXC>
struct WalkPlanOf {
  static constexpr WalkPlan<N, GS> v = make_walk_plan<N, GS, PXC>();
};

}  // namespace gse
