"""GSE-u and GSE-r on the GPU against the FP64 oracle.

FP64 (bar 1e-10): energies, reference forces, potential grids and the
intermediate grids of the staged calls, for the reference's pair and a
water box, cubic and anisotropic; GSE-u with the reference's testPair
parameters (N = 4, b = 1.168) and with convolution radii longer than the
grid (b = 2).  GSE-r solves the reference's HERMITIAN system exactly; the
reference relaxes 200 sweeps towards it, and on its pair at M = 16 the two
agree to the relaxation's own residual (energy 1.4e-8 relative, observed
on the CPU with the reference compiled), which this test bounds at 1e-7.
FP32: Erms <= 1e-5 against the FP64 oracle.
"""
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


CASES = [
    ("pair", lambda: systems.pair(1.0, 16.0, orientation="diagonal"), (16, 16, 16), (16.0,) * 3, 2.121320344),
    ("water", lambda: systems.water_box(300, 24.0, seed=14), (24, 24, 24), (24.0,) * 3, 2.12),
    ("aniso", lambda: systems.water_box(100, 16.0, seed=9) * np.array([1.0, 0.8, 1.2, 1.0]), (16, 12, 20), (16.0, 12.8, 19.2), 2.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gser_fp64(case):
    name, make, dims, box, s = case
    x = make()
    p = orc.Params(dims, box, s)
    e, f, rho, rhoS, phi = orc.gser_step(p, x, nsweep=0)
    g = api.CpuGSEr(F64, F64, *dims, s, 0.4, 3.0, 3.0, *box)
    xd = dev(x)
    g.spreadCharge1(len(x), xd)
    g.spreadCharge2()
    g.solvePoisson()
    E = g.calculateEnergy()
    fo = torch.zeros((3, len(x)), dtype=F64, device="cuda")
    g.interpolateForce(len(x), xd, fo[0], fo[1], fo[2])
    assert abs(E - e) <= TOL * abs(e), name
    assert rel(host(fo).T, f) <= TOL, name
    assert rel(host(g.getRho()), rho) <= TOL
    assert rel(host(g.getRhoS()), rhoS) <= TOL
    assert rel(host(g.getPhi()), phi) <= TOL
    o = g.step(xd, force=True)
    assert abs(float(host(o["energy"])[0]) - e) <= TOL * abs(e)
    assert rel(host(o["force"]).T, f) <= TOL


def test_gser_against_the_reference_relaxation():
    x = systems.pair(1.0, 16.0, orientation="diagonal")
    p = orc.Params(16, 16.0, 2.121320344)
    e200, _, _, _, phi200 = orc.gser_step(p, x, nsweep=200)   # bit-identical to CpuGSEr
    g = api.CpuGSEr(F64, F64, 16, 16, 16, p.sigma, 0.4, 3.0, 3.0, 16.0, 16.0, 16.0)
    o = g.step(dev(x), force=False)
    E = float(host(o["energy"])[0])
    print(f"GSE-r pair M=16: exact {E:.12f}, 200 sweeps {e200:.12f}", flush=True)
    assert abs(E - e200) <= 1e-7 * abs(e200)
    phi = host(g.getPhi())
    d = (phi200 - phi200.mean()) - (phi - phi.mean())
    assert np.max(np.abs(d)) <= 1e-6 * np.max(np.abs(phi))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("N,b", [(4, 1.168), (3, 2.0)], ids=["N4_b1.168", "N3_b2"])
def test_gseu_fp64(case, N, b):
    name, make, dims, box, s = case
    x = make()
    p = orc.Params(dims, box, s)
    e, f, phi = orc.gseu_step(p, N, b, x)
    g = api.CpuGSEu(F64, F64, N, b, *dims, s, 0.4, 3.0, 3.0, *box)
    xd = dev(x)
    g.spreadCharge1(len(x), xd)
    g.solvePoisson()
    E = g.calculateEnergy()
    fo = torch.zeros((3, len(x)), dtype=F64, device="cuda")
    g.interpolateForce(len(x), xd, fo[0], fo[1], fo[2])
    assert abs(E - e) <= TOL * abs(e), (name, E, e)
    assert rel(host(fo).T, f) <= TOL, name
    assert rel(host(g.getPhi()), phi) <= TOL
    o = g.step(xd, force=True)
    assert abs(float(host(o["energy"])[0]) - e) <= TOL * abs(e)
    assert rel(host(o["force"]).T, f) <= TOL


def test_gseu_gser_fp32_water():
    xyzq, L, M, s = systems.make("water_3k_M32")
    x32 = dev(xyzq, F32)
    xr = x32.double().cpu().numpy()
    p = orc.Params(M, L, s)
    e, f, _, _, _ = orc.gser_step(p, xr)
    o = api.CpuGSEr(F32, F32, M, M, M, s, 0.4, 3.0, 3.0, L, L, L).step(x32, force=True)
    assert orc.erms(f, host(o["force"]).T) <= 1e-5
    assert abs(float(host(o["energy"])[0]) - e) <= 1e-5 * abs(e)
    e, f, _ = orc.gseu_step(p, 4, 1.168, xr)
    o = api.CpuGSEu(F32, F32, 4, 1.168, M, M, M, s, 0.4, 3.0, 3.0, L, L, L).step(x32, force=True)
    assert orc.erms(f, host(o["force"]).T) <= 1e-5
    assert abs(float(host(o["energy"])[0]) - e) <= 1e-5 * abs(e)


def test_fft_solver_hermitian_kind():
    """CpuFFTSolver with the HERMITIAN kind is the exact solve of that system."""
    x = systems.random_alternating(30, 18.0, seed=5)
    p = orc.Params((18, 16, 20), (18.0, 16.0, 20.0), 2.0)
    _, _, _, rhoS, _ = orc.gser_step(p, x)
    ref = orc.mehrstellen_solve(p.dims, p.box, rhoS)
    s = api.CpuFFTSolver(api.CpuGreensFuncHermitian(), F64, *p.dims)
    phi = torch.zeros((20, 16, 18), dtype=F64, device="cuda")
    s.run(phi, dev(rhoS), 0.0, *p.box)
    assert rel(host(phi), ref) <= TOL
