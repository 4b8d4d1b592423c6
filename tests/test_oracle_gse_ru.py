"""CPU checks of the oracle's GSE-u and GSE-r restatements.

  * GSE-u (CpuGSEu.cpp) and GSE-r with the reference's 200 red-black sweeps
    (CpuGSEr.cpp + CpuMultiGridSolver.cpp) are bit-identical to the
    reference's own classes compiled in oracle/_ref;
  * the k-space solve of the HERMITIAN system (the form the GPU uses) is a
    fixed point of the reference's relaxation and the limit of its sweeps
    when the source is consistent (zero grid sum), on cubic and anisotropic
    grids; this pins the stencil symbols A^ and B^;
  * near k = 0 the symbol reduces to the Coulomb kernel 4 pi/(k^2 XYZ).
"""
import math
import os

import numpy as np
import pytest

from gse_amd import systems


@pytest.fixture
def in_tmp(tmp_path):
    """The reference classes write text files into the working directory."""
    old = os.getcwd()
    os.chdir(tmp_path)
    yield tmp_path
    os.chdir(old)


@pytest.mark.parametrize("case", [("pair", 16, 16.0, 2.121320344),
                                  ("water", 20, 19.0, 2.12)], ids=lambda c: c[0])
def test_gseu_and_gser_bit_identical_to_reference(orc, ref, in_tmp, case):
    name, M, L, s = case
    x = systems.pair(1.0, L, orientation="diagonal") if name == "pair" else \
        systems.water_box(60, L, seed=4)
    p = orc.Params(M, L, s)
    e1, f1, phi1 = orc.gseu_step(p, 4, 1.168, x)
    e2, f2, phi2 = orc.ref_gseu_step(p, 4, 1.168, x)
    assert e1 == e2 and np.array_equal(f1, f2) and np.array_equal(phi1, phi2)
    e3, f3, _, rs3, ph3 = orc.gser_step(p, x, nsweep=200)
    e4, f4, rs4, ph4 = orc.ref_gser_step(p, x)
    assert e3 == e4 and np.array_equal(f3, f4)
    assert np.array_equal(rs3, rs4) and np.array_equal(ph3, ph4)


@pytest.mark.parametrize("dims,box", [((16, 16, 16), (16.0,) * 3),
                                      ((16, 12, 20), (16.0, 13.0, 19.0))])
def test_exact_hermitian_solve_is_the_relaxation_limit(orc, dims, box):
    x = systems.random_alternating(24, min(box), seed=3)
    p = orc.Params(dims, box, 2.0)
    _, _, _, rs, _ = orc.gser_step(p, x)
    rs = rs - rs.mean()   # a consistent source: zero grid sum
    ex = orc.mehrstellen_solve(dims, box, rs)
    assert abs(ex.mean()) <= 1e-13 * np.max(np.abs(ex))
    one = orc.rbgs_hermitian(dims, box, -4 * np.pi * rs, ex, 1)
    assert np.max(np.abs(one - ex)) <= 1e-13 * np.max(np.abs(ex))
    it = orc.rbgs_hermitian(dims, box, -4 * np.pi * rs, np.zeros_like(rs), 800)
    assert np.max(np.abs((it - it.mean()) - ex)) <= 1e-11 * np.max(np.abs(ex))


def test_hermitian_symbol_tends_to_coulomb(orc):
    X = Y = Z = 64
    L = 64.0
    for m in ((1, 0, 0), (0, 1, 1), (1, 1, 1)):
        g = orc.lib().orc_mehrstellen_g
        g.restype = __import__("ctypes").c_double
        d = __import__("ctypes").c_double
        val = g(X, Y, Z, d(L), d(L), d(L), *m)
        k2 = sum((2 * math.pi * mi / L) ** 2 for mi in m)
        coul = 4 * math.pi / (k2 * X * Y * Z)
        assert abs(val - coul) / coul < 3e-3   # O((kh)^4) stencil error at kh ~ 0.1
    assert g(X, Y, Z, d(L), d(L), d(L), 0, 0, 0) == 0.0
