"""The SPME reciprocal comparator (CudaPMERecip) and the Ewald direct-space
sum (CpuEwaldDirect) on the GPU against the FP64 oracle.

SPME, FP64 (bar 1e-10 relative): charge grid, energy, virial and forces for
orders 4, 6 and 8 on the reference's pair (C1, "SPME order 4"), a water box,
an anisotropic box, an odd-sized grid and negative coordinates; the staged
calls (spread_charge, r2c_fft, scalar_sum, c2r_fft, gather_force with the
reference's gradient sign) and the self energy; the grid bit-identical on
reruns and under a permutation of the particles.  FP32: Erms <= 1e-5 against
the FP64 oracle on the 48k water box.

Ewald direct, FP64 (1e-10): energy and dE/dr with and without the
reference's exclusion lists, a small box (all-pairs path), an anisotropic
box and coordinates outside the box; reruns bit-identical.  A test48K-style
total (GSE-k + Ewald direct against Ewald reciprocal + Ewald direct)
closes the loop at the approximation level of the method.
"""
import math

import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

F64, F32 = torch.float64, torch.float32
TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b)))


def dev(a, dt=F64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.detach().double().cpu().numpy()


SPME_CASES = [
    ("pair_diag_o4", lambda: systems.pair(1.0, 20.0, orientation="diagonal"), (20, 20, 20), (20.0,) * 3, 4, 2.121320344),
    ("pair_seed1_o4", lambda: systems.pair(1.0, 20.0, seed=1), (20, 20, 20), (20.0,) * 3, 4, 2.121320344),
    ("water3k_o4", lambda: systems.water_box(1000, 31.0, seed=3), (32, 32, 32), (31.0,) * 3, 4, 2.12),
    ("water3k_o6", lambda: systems.water_box(1000, 31.0, seed=3), (32, 32, 32), (31.0,) * 3, 6, 2.12),
    ("water3k_o8", lambda: systems.water_box(1000, 31.0, seed=3), (32, 32, 32), (31.0,) * 3, 8, 2.12),
    ("aniso_o6", lambda: systems.water_box(150, 18.0, seed=7) * np.array([1.0, 0.9, 1.2, 1.0]), (18, 16, 24), (18.0, 16.2, 21.6), 6, 2.0),
    ("odd_o4", lambda: systems.random_alternating(60, 21.0, seed=9), (21, 20, 25), (21.0, 20.0, 25.0), 4, 2.0),
    ("negative_o8", lambda: systems.random_alternating(40, 16.0, seed=6) - np.array([8.0, 3.0, 15.0, 0.0]), (16, 16, 16), (16.0,) * 3, 8, 2.0),
]


@pytest.mark.parametrize("case", SPME_CASES, ids=[c[0] for c in SPME_CASES])
def test_spme_step_fp64(case):
    name, make, dims, box, order, sigma = case
    x = make()
    kappa = 1.0 / (sigma * math.sqrt(2.0))
    e, vir, f, Q = orc.pme_step(dims, order, x, box, kappa, want_grid=True)
    pm = api.CudaPMERecip(F64, F64, *dims, order)
    o = pm.step(dev(x), box, kappa)
    assert abs(float(host(o["energy"])[0]) - e) <= TOL * abs(e), name
    assert rel(host(o["virial"]).reshape(3, 3), vir) <= TOL, name
    assert rel(host(o["force"]).T, f) <= TOL, name
    # the staged calls of CudaPMERecip.h:155-167 in gpu_gse.cu:61-73 order
    recip = [1.0 / box[0], 0, 0, 0, 1.0 / box[1], 0, 0, 0, 1.0 / box[2]]
    pm.spread_charge(dev(x), len(x), recip)
    assert rel(host(pm.grid()), Q) <= TOL, name
    pm.r2c_fft()
    s = pm.scalar_sum(recip, kappa, True, True)
    pm.c2r_fft()
    g = pm.gather_force(dev(x), len(x), recip, stride=len(x) + 3)
    assert abs(float(host(s["energy"])[0]) - e) <= TOL * abs(e)
    assert rel(host(s["virial"]).reshape(3, 3), vir) <= TOL
    assert rel(host(g)[:, :len(x)].T, -f) <= TOL
    self_e = float(host(pm.calc_self_energy(dev(x), len(x), kappa))[0])
    exact = -kappa * 332.0716 / math.sqrt(math.pi) * float(np.sum(x[:, 3] ** 2))
    assert abs(self_e - exact) <= 1e-13 * abs(exact)


def test_spme_grid_deterministic_and_order_independent():
    x = systems.water_box(1000, 31.0, seed=3)
    pm = api.CudaPMERecip(F32, F32, 32, 32, 32, 4)
    recip = (31.0, 31.0, 31.0)
    pm.spread_charge(dev(x, F32), len(x), recip)
    g1 = pm.grid().clone()
    pm.spread_charge(dev(x, F32), len(x), recip)
    assert torch.equal(pm.grid(), g1)
    perm = np.random.default_rng(4).permutation(len(x))
    pm.spread_charge(dev(x[perm], F32), len(x), recip)
    assert torch.equal(pm.grid(), g1)


def test_spme_fp32_water48k_force_rms():
    xyzq, L, M, sigma = systems.make("water_48k_M80")
    x32 = dev(xyzq, F32)
    xr = x32.double().cpu().numpy()
    kappa = 1.0 / (sigma * math.sqrt(2.0))
    for order in (4, 8):
        e, _, f = orc.pme_step((M, M, M), order, xr, (L,) * 3, kappa)
        pm = api.CudaPMERecip(F32, F32, M, M, M, order)
        o = pm.step(x32, (L, L, L), kappa)
        er = orc.erms(f, host(o["force"]).T)
        de = abs(float(host(o["energy"])[0]) - e) / abs(e)
        print(f"SPME{order} FP32 48k: Erms {er:.2e} dE/E {de:.2e}", flush=True)
        assert er <= 1e-5
        assert de <= 1e-5


def test_spme_bad_arguments():
    with pytest.raises(api._lib.GSEError):
        api.CudaPMERecip(F64, F64, 16, 16, 16, 5)
    with pytest.raises(api._lib.GSEError):
        api.CudaPMERecip(F64, F64, 3, 16, 16, 4)
    pm = api.CudaPMERecip(F64, F64, 16, 16, 16, 4)
    x = dev(systems.random_alternating(10, 16.0, seed=1))
    with pytest.raises(api._lib.GSEError):  # triclinic
        pm.spread_charge(x, 10, [1 / 16.0, 0.01, 0, 0, 1 / 16.0, 0, 0, 0, 1 / 16.0])


def water_excl(n):
    ib, nb = [], []
    for j in range(n):
        t = j // 3
        for i in range(3 * t, 3 * t + 3):
            if i > j:
                nb.append(i + 1)
        ib.append(len(nb))
    return ib, nb


EW_CASES = [
    ("water3k", lambda: systems.water_box(1000, 31.0, seed=3), (31.0,) * 3, 9.0, 2.12, True),
    ("water3k_noexcl", lambda: systems.water_box(1000, 31.0, seed=3), (31.0,) * 3, 9.0, 2.12, False),
    ("small_box_all_pairs", lambda: systems.random_alternating(200, 20.0, seed=5), (20.0,) * 3, 9.0, 2.0, False),
    ("aniso", lambda: systems.water_box(400, 24.0, seed=7) * np.array([1.0, 1.3, 0.8, 1.0]), (24.0, 31.2, 19.2), 6.0, 1.8, True),
    ("outside_box", lambda: systems.water_box(500, 25.0, seed=8, wrap=False) - np.array([3.0, -2.0, 1.0, 0.0]), (25.0,) * 3, 7.5, 2.12, True),
]


@pytest.mark.parametrize("case", EW_CASES, ids=[c[0] for c in EW_CASES])
def test_ewald_direct_fp64(case):
    name, make, box, rcut, sigma, excl = case
    x = make()
    n = len(x)
    ib, nb = water_excl(n) if excl else (None, None)
    e, g = orc.ewald_direct(sigma, rcut, box, x, ib, nb)
    d = api.CpuEwaldDirect(sigma, rcut, *box)
    if excl:
        d.setupExclusions(n, ib, nb)
    f = torch.zeros((3, n), dtype=F64, device="cuda")
    eg = d.calcForceEnergy(n, dev(x), f[0], f[1], f[2])
    assert abs(eg - e) <= TOL * abs(e), (name, eg, e)
    assert rel(host(f).T, g) <= TOL, name
    f2 = torch.zeros_like(f)
    eg2 = d.calcForceEnergy(n, dev(x), f2[0], f2[1], f2[2])
    assert eg2 == eg and torch.equal(f2, f)


def test_total_electrostatics_like_test48k():
    """cpu_gse.cpp:332-389 adds the direct-space sum to each reciprocal method
    and compares with Ewald.  On the 3k water box, GSE-k (this library) plus
    the Ewald direct sum (this library) against the Ewald reciprocal sum
    (oracle, kcut 14) plus the direct sum: energies agree to the method's
    accuracy (observed 2.7e-9 relative in FP64), forces to Erms 1e-4."""
    xyzq, L, M, sigma = systems.make("water_3k_M32")
    x = dev(xyzq)
    n = len(xyzq)
    g = api.CpuGSEk(F64, F64, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    o = g.step(x, force=True, potential=False)
    d = api.CpuEwaldDirect(sigma, 9.0, L, L, L)
    gd = torch.zeros((3, n), dtype=F64, device="cuda")
    ed = d.calcForceEnergy(n, x, gd[0], gd[1], gd[2])
    er, fr = orc.ewald_recip(sigma, 14, (L,) * 3, xyzq)
    tot_gse = float(host(o["energy"])[0]) + ed
    tot_ew = er + ed
    # CpuEwaldDirect returns dE/dr, the reciprocal methods forces: -dE/dr
    f_gse = host(o["force"]).T - host(gd).T
    f_ew = fr - host(gd).T
    print(f"total E: GSE-k+direct {tot_gse:.8f} Ewald {tot_ew:.8f}", flush=True)
    assert abs(tot_gse - tot_ew) / abs(tot_ew) <= 1e-7
    assert orc.erms(f_ew, f_gse) <= 1e-4


@pytest.mark.parametrize("case", [
    ("pair_kcut32", lambda: systems.pair(1.0, 20.0, orientation="diagonal"), (20.0,) * 3, 32, 2.121320344),
    ("box300_kcut10", lambda: systems.water_box(100, 18.0, seed=12), (18.0, 18.0, 18.0), 10, 2.12),
    ("aniso_kcut8", lambda: systems.random_alternating(50, 15.0, seed=3) * np.array([1.0, 1.2, 0.8, 1.0]), (15.0, 18.0, 12.0), 8, 1.7),
], ids=lambda c: c[0])
def test_ewald_recip_fp64(case):
    """CpuEwaldRecip on the GPU against the oracle's restatement (itself
    pinned against the reference's compiled class): energy and forces."""
    name, make, box, kcut, sigma = case
    x = make()
    n = len(x)
    e, f = orc.ewald_recip(sigma, kcut, box, x)
    er = api.CpuEwaldRecip(sigma, kcut, *box)
    fg = torch.zeros((3, n), dtype=F64, device="cuda")
    eg = er.calcForceEnergy(n, dev(x), fg[0], fg[1], fg[2])
    assert abs(eg - e) <= TOL * abs(e), (name, eg, e)
    assert rel(host(fg).T, f) <= TOL, name
    assert er.calcForceEnergy(n, dev(x)) == eg
