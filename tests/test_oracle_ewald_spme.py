"""CPU checks of the oracle's Ewald direct-space sum and SPME reciprocal path.

Ewald direct (CpuEwaldDirect.cpp:45-131):
  * bit-identical to the reference's own compiled class (oracle/_ref), with
    and without an exclusion list;
  * a single pair against the closed form q1 q2 erfc(r/(sqrt2 sigma))/r and
    its derivative, plus the self term -ccelec sum q^2/sqrt(2 pi sigma^2);
  * gradient = central finite differences of the energy; the minimum image.
SPME (cuda/CudaPMERecip.cu, restated in oracle/spme_oracle.c):
  * B-spline weights equal the cardinal B-spline M_n(w + n - 1 - j) from its
    closed form, sum to one, and their derivatives are those of M_n;
  * moduli: prefac[0] = 1 and prefac[k] = prefac[M - k];
  * the energy converges to the Ewald reciprocal sum (orc_ewald_recip, pinned
    against the reference's CpuEwaldRecip) as order and grid grow;
  * forces = -dE/dr and virial = -dE/d eps by central finite differences;
  * the energy is unchanged by a translation of every charge by one grid
    spacing.
"""
import math

import numpy as np
import pytest

from gse_amd import systems

CC = 332.0716


def water_excl(n):
    """Exclusions of each rigid triplet's internal pairs, reference format."""
    ib, nb = [], []
    for j in range(n):
        t = j // 3
        for i in range(3 * t, 3 * t + 3):
            if i > j:
                nb.append(i + 1)
        ib.append(len(nb))
    return ib, nb


@pytest.mark.parametrize("excl", [False, True])
def test_ewald_direct_bit_identical_to_reference(orc, ref, excl):
    x = systems.water_box(200, 22.0, seed=21)
    ib, nb = water_excl(len(x)) if excl else (None, None)
    e1, g1 = orc.ewald_direct(2.12, 9.0, (22.0,) * 3, x, ib, nb)
    e2, g2 = orc.ewald_direct(2.12, 9.0, (22.0,) * 3, x, ib, nb, reference=True)
    assert e1 == e2
    assert np.array_equal(g1, g2)


def test_ewald_direct_pair_closed_form(orc):
    s, r = 2.12, 3.7
    x = np.array([[1.0, 2.0, 3.0, 0.8], [1.0 + r, 2.0, 3.0, -0.6]])
    e, g = orc.ewald_direct(s, 9.0, (30.0,) * 3, x)
    a = 1.0 / (math.sqrt(2.0) * s)
    qq = 0.8 * -0.6 * CC
    pair = qq * math.erfc(a * r) / r
    self_e = -CC * (0.8 ** 2 + 0.6 ** 2) / math.sqrt(2.0 * math.pi * s * s)
    assert abs(e - (pair + self_e)) <= 1e-13 * abs(pair + self_e)
    dpair = qq * (-2.0 * a / math.sqrt(math.pi) * math.exp(-(a * r) ** 2) / r
                  - math.erfc(a * r) / r ** 2)
    assert abs(g[1, 0] - dpair) <= 1e-12 * abs(dpair)   # dE/dx of atom 1
    assert abs(g[0, 0] + dpair) <= 1e-12 * abs(dpair)
    assert np.all(g[:, 1:] == 0.0)
    # beyond the cutoff only the self term remains
    e2, _ = orc.ewald_direct(s, 3.0, (30.0,) * 3, x)
    assert abs(e2 - self_e) <= 1e-14 * abs(self_e)


def test_ewald_direct_minimum_image_and_fd(orc):
    L = 18.0
    x = systems.random_alternating(30, L, seed=8)
    x[0, :3] = [0.3, 0.2, 17.9]   # a pair across the periodic boundary
    x[1, :3] = [17.8, 0.1, 0.4]
    e, g = orc.ewald_direct(2.0, 8.5, (L,) * 3, x)
    h = 1e-5
    for i in (0, 1, 7):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd = (orc.ewald_direct(2.0, 8.5, (L,) * 3, xp)[0]
                  - orc.ewald_direct(2.0, 8.5, (L,) * 3, xm)[0]) / (2 * h)
            assert abs(fd - g[i, c]) <= 1e-6 * max(1.0, abs(g[i, c])), (i, c, fd, g[i, c])


def _bspline(n, x):
    """Cardinal B-spline M_n(x) from its closed form."""
    s = 0.0
    for k in range(n + 1):
        t = x - k
        if t > 0:
            s += (-1) ** k * math.comb(n, k) * t ** (n - 1)
    return s / math.factorial(n - 1)


@pytest.mark.parametrize("order", [4, 6, 8])
def test_pme_theta_is_the_cardinal_bspline(orc, order):
    for w in (0.0, 0.137, 0.5, 0.731, 0.999):
        th, dth = orc.pme_theta(order, w)
        exact = [_bspline(order, w + order - 1 - j) for j in range(order)]
        dexact = [_bspline(order - 1, w + order - 1 - j) - _bspline(order - 1, w + order - 2 - j)
                  for j in range(order)]
        assert np.max(np.abs(th - exact)) <= 1e-12  # closed form loses ~1e-14 to cancellation
        assert np.max(np.abs(dth - dexact)) <= 1e-12
        assert abs(th.sum() - 1.0) <= 1e-14
        assert abs(dth.sum()) <= 1e-14


@pytest.mark.parametrize("order", [4, 6, 8])
def test_pme_moduli(orc, order):
    for M in (20, 24, 31):
        p = orc.pme_prefac(order, M)
        assert abs(p[0] - 1.0) <= 1e-14
        assert np.allclose(p[1:], p[1:][::-1], rtol=1e-12)
        assert np.all(p >= 1.0 - 1e-14)


def test_pme_energy_converges_to_ewald(orc):
    L = 20.0
    x = systems.water_box(100, L, seed=5)
    kappa = 1.0 / (2.12 * math.sqrt(2.0))
    er, _ = orc.ewald_recip(2.12, 20, (L,) * 3, x)
    errs = {}
    for order, M in ((4, 20), (4, 48), (6, 32), (8, 32), (8, 48)):
        e, _, _ = orc.pme_step((M, M, M), order, x, (L,) * 3, kappa)
        errs[(order, M)] = abs(e - er) / abs(er)
    assert errs[(4, 48)] < errs[(4, 20)] / 10
    assert errs[(8, 32)] < errs[(6, 32)] / 10
    assert errs[(8, 48)] < 1e-8


def test_pme_forces_and_virial_are_energy_derivatives(orc):
    L, M = 16.0, 20
    x = systems.random_alternating(24, L, seed=11)
    kappa = 1.0 / (2.0 * math.sqrt(2.0))
    for order in (4, 6):
        e, vir, f = orc.pme_step((M, M, M), order, x, (L,) * 3, kappa)
        h = 1e-5
        for i in (0, 5, 13):
            for c in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, c] += h
                xm[i, c] -= h
                fd = -(orc.pme_step((M, M, M), order, xp, (L,) * 3, kappa)[0]
                       - orc.pme_step((M, M, M), order, xm, (L,) * 3, kappa)[0]) / (2 * h)
                assert abs(fd - f[i, c]) <= 1e-6 * np.max(np.abs(f)), (order, i, c)
        # W_aa = -dE/d eps_aa: stretch axis a (box and positions)
        eps = 1e-6
        for a in range(3):
            sp, sm = np.ones(3), np.ones(3)
            sp[a] += eps
            sm[a] -= eps
            xp, xm = x.copy(), x.copy()
            xp[:, :3] *= sp
            xm[:, :3] *= sm
            ep = orc.pme_step((M, M, M), order, xp, tuple(L * sp), kappa)[0]
            em = orc.pme_step((M, M, M), order, xm, tuple(L * sm), kappa)[0]
            assert abs(-(ep - em) / (2 * eps) - vir[a, a]) <= 1e-6 * abs(e), (order, a)
        assert np.allclose(vir, vir.T, rtol=0, atol=1e-12 * abs(e))


def test_pme_energy_invariant_under_grid_translation(orc):
    L, M = 16.0, 16
    x = systems.random_alternating(20, L, seed=2)
    kappa = 0.3
    e0, _, f0 = orc.pme_step((M, M, M), 4, x, (L,) * 3, kappa)
    y = x.copy()
    y[:, :3] += np.array([1.0, 2.0, -3.0]) * (L / M)
    e1, _, f1 = orc.pme_step((M, M, M), 4, y, (L,) * 3, kappa)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    assert np.max(np.abs(f1 - f0)) <= 1e-11 * np.max(np.abs(f0))
