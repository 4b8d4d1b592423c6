"""Writes tests/golden/ref_vectors.npz from the reference's own compiled CPU
classes (oracle/_ref/libgse_ref.so, built by `make -C oracle ref` from the
reference sources).  Only the numbers are committed.

Inputs are regenerated in the test from the seeds stored here, so the
fixture holds just the reference outputs:

  pair_rho       CpuGaussCharge<double,double>::spreadChargeToGrid, the
                 two-charge system of cpu_gse.cpp:689-698 (r=1, L=20, M=20,
                 sigma1 = 2.121320344*sqrt(0.4), rcut = 3*sqrt(2)*sigma1)
  pair_force     interpolateForce on the same system with a seeded phi grid
  rand_rho       spreadChargeToGrid(normalize=true) on 25 seeded charges,
                 12x10x14 grid
  greens_gse     CpuGreensFuncGSE<double>::mul on a seeded 9x6x7 spectrum
  greens_g2      CpuGreensFuncG2<double>::mul on the same spectrum
  conv_x         CpuConvolution<double>::run (X pass) on a seeded 16x8x6 grid
  ongrid         spreadChargeOnGrid on a seeded 12x10x14 grid

Run:  python tests/golden/make_golden.py
"""
import ctypes
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests.golden import inputs  # noqa: E402

dp = ctypes.POINTER(ctypes.c_double)
D = ctypes.c_double


def P(a):
    return a.ctypes.data_as(dp)


def main():
    oracle.build(ref=True)
    R = oracle.ref_lib()
    out = {}

    d = inputs.pair_case()
    X, Y, Z = d["dims"]
    rho = np.zeros((Z, Y, X))
    a = np.ascontiguousarray(d["xyzq"])
    R.ref_spread_dd(X, Y, Z, D(d["sigma"]), D(d["rcut"]), 2, P(a), *map(D, d["box"]), P(rho), 0)
    out["pair_rho"] = rho
    f = np.zeros((3, 2))
    phi = np.ascontiguousarray(d["phi"])
    R.ref_interpolate_dd(X, Y, Z, D(d["sigma"]), D(d["rcut"]), 2, P(a), *map(D, d["box"]), P(phi),
                         P(f[0]), P(f[1]), P(f[2]))
    out["pair_force"] = f.T.copy()

    d = inputs.rand_case()
    X, Y, Z = d["dims"]
    rho = np.zeros((Z, Y, X))
    a = np.ascontiguousarray(d["xyzq"])
    R.ref_spread_dd(X, Y, Z, D(d["sigma"]), D(d["rcut"]), len(a), P(a), *map(D, d["box"]), P(rho), 1)
    out["rand_rho"] = rho

    d = inputs.greens_case()
    X, Y, Z = d["dims"]
    for kind, key in ((0, "greens_gse"), (1, "greens_g2")):
        s = np.ascontiguousarray(d["spec"].copy())
        R.ref_greens_mul_d(kind, X, Y, Z, *map(D, d["box"]), D(d["sigma"]), s.ctypes.data_as(dp))
        out[key] = s

    d = inputs.conv_case()
    X, Y, Z = d["dims"]
    g = np.ascontiguousarray(d["grid"])
    o = np.zeros_like(g)
    k = np.ascontiguousarray(d["kx"])
    R.ref_conv_run_d(X, Y, Z, d["R"], d["R"], d["R"], P(k), P(k), P(k), P(g), P(o))
    out["conv_x"] = o

    d = inputs.ongrid_case()
    X, Y, Z = d["dims"]
    g = np.ascontiguousarray(d["grid"])
    o = np.zeros_like(g)
    R.ref_spread_on_grid_dd(X, Y, Z, D(d["sigma"]), D(d["rcut"]), *map(D, d["box"]), P(g), P(o))
    out["ongrid"] = o

    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
