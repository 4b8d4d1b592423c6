"""The gather's compile-time walk plan (gse_amd/csrc/walk_plan.h), checked on
the host: the column order is a permutation whose 32- or 16-lane slots hit
distinct LDS words modulo 64 (13^3 cubes over 20 x 20 brick planes), and
every z step a slot marks inside (or
outside) the sphere is inside (outside) for every charge position the
kernel can see, by brute force over charge offsets within half a cell of
the cube centre and cutoffs rcut/h in (n-1, n] on each axis (anisotropic
spacings included).  The sphere test is the gather's r^2 < rcut^2
(CpuGaussCharge.cpp:157)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "gse_amd", "csrc")

CHECK = r"""
#include <cstdio>
#include <random>
#include "walk_plan.h"
using namespace gse;

template <int N, int GS, int PXC, bool DISTINCT>
int check() {
  constexpr WalkPlan<N, GS> p = make_walk_plan<N, GS, PXC>();
  constexpr int n = N / 2, NC = N * N, KS = WalkPlan<N, GS>::KS;
  int seen[NC] = {};
  for (int c = 0; c < NC; c++) {
    if (p.col[c] < 0 || p.col[c] >= NC || seen[p.col[c]]++) { std::printf("not a permutation\n"); return 1; }
  }
  for (int k = 0; k < KS && DISTINCT; k++) {
    bool bank[64] = {};
    for (int l = 0; l < GS && GS * k + l < NC; l++) {
      const int c = p.col[GS * k + l];
      const int r = (c % N + PXC * (c / N)) % 64;
      if (bank[r]) { std::printf("slot %d: LDS word %d twice\n", k, r); return 1; }
      bank[r] = true;
    }
  }
  int nin = 0, nout = 0;
  for (int k = 0; k < KS; k++)
    for (int t = 0; t < N; t++) nin += p.cls[k][t] == 1, nout += p.cls[k][t] == 2;
  if (N == 13 && PXC > 0 && (nin == 0 || nout == 0)) { std::printf("trivial plan\n"); return 1; }
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> uf(-0.5, 0.5), ur(0.0, 1.0);
  long long tested = 0;
  for (int trial = 0; trial < 4000; trial++) {
    // spacings h_a with rcut/h_a in (n-1, n] on every axis
    double h[3], f[3];
    const double rc = 1.0;
    for (int a = 0; a < 3; a++) {
      const int mode = trial % 4;
      const double q = mode == 0 ? 1e-9 : mode == 1 ? 1.0 : ur(rng);  // position in (n-1, n]
      h[a] = rc / (n - 1 + (mode == 2 && a == 1 ? 1.0 : q));
      if (h[a] <= 0 || rc / h[a] <= n - 1 || rc / h[a] > n) h[a] = rc / (n - 0.5);
      f[a] = trial < 8 ? ((trial >> (a % 3)) & 1 ? 0.5 : -0.5) : uf(rng);
    }
    for (int k = 0; k < KS; k++)
      for (int l = 0; l < GS && GS * k + l < NC; l++) {
        const int c = p.col[GS * k + l];
        const int o[2] = {c % N - n, c / N - n};
        for (int t = 0; t < N; t++) {
          const int cl = p.cls[k][t];
          if (cl == 0) continue;
          const double dx = (o[0] - f[0]) * h[0], dy = (o[1] - f[1]) * h[1],
                       dz = (t - n - f[2]) * h[2];
          const bool in = dx * dx + dy * dy + dz * dz < rc * rc;
          tested++;
          if ((cl == 1) != in) {
            std::printf("N=%d GS=%d slot %d col (%d,%d) z %d class %d but %s\n", N, GS, k, o[0],
                        o[1], t - n, cl, in ? "inside" : "outside");
            return 1;
          }
        }
      }
  }
  std::printf("N=%d GS=%d PXC=%d: %d inside, %d outside of %d slot steps; %lld point checks\n", N, GS,
              PXC, nin, nout, KS * N, tested);
  return 0;
}

int main() {
  // 13^3 over 20 x 20 brick planes: every slot on distinct LDS words
  return check<13, 32, 20, true>() | check<13, 16, 20, true>() | check<15, 32, 24, false>() |
         check<13, 32, 0, false>();
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_walk_plan_classes_hold_for_every_charge_position():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "check.cpp")
        exe = os.path.join(d, "check")
        with open(src, "w") as fh:
            fh.write(CHECK)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-o", exe, src], check=True,
                       capture_output=True, text=True, timeout=120)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
