"""The 1,000,002-charge water box (333,334 triplets, L = 215.3, M = 216,
sigma = 2.12), FP32 forward and adjoint.

The FP64 oracle is run as the composition of its pinned pieces (spread,
FFT solve, energy) on the same float32-rounded particles, and its gathers
on a 4,000-particle sample, so the test finishes in tens of seconds.
Size-independent properties are checked on the whole system: bit-identical
grids on reruns and under particle permutation, exact permutation
equivariance of the per-particle outputs, and Parseval (k-space energy ==
real-space sum rho*phi).
"""
import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

F32 = torch.float32
NSAMPLE = 4000


@pytest.fixture(scope="module")
def big():
    xyzq, L, M, sigma = systems.make("water_1M_M216")
    x32 = torch.from_numpy(xyzq).to("cuda", F32).contiguous()
    xr = x32.double().cpu().numpy()
    p = orc.Params(M, L, sigma)
    print("oracle: spread + solve of 1,000,002 charges at M=216", flush=True)
    rho = orc.spread(p.dims, p.sigma1, p.rcut1, xr, p.box)
    phi = orc.fft_solve(0, p.dims, p.sigma2, p.box, rho)
    E = orc.energy(p.dims, p.ccelec, p.box, rho, phi)
    idx = np.sort(np.random.default_rng(5).choice(len(xr), NSAMPLE, replace=False))
    fs = orc.interpolate(p.dims, p.sigma1, p.rcut1, xr[idx], p.box, phi) * p.ccelec
    pot, fc, _ = orc.gather_cube(p.dims, p.sigma1, p.rcut1, xr[idx], p.box, phi)
    print("oracle done", flush=True)
    g = api.CpuGSEk(F32, F32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    return dict(x=x32, p=p, E=E, idx=idx, fs=fs, pot=pot * p.ccelec, fc=fc * p.ccelec, g=g)


def test_forward_fp32_vs_oracle(big):
    g, idx = big["g"], big["idx"]
    o = g.step(big["x"], force=True, potential=True, cube=True)
    torch.cuda.synchronize()
    e = float(o["energy"].cpu()[0])
    f = o["force"].double().cpu().numpy().T[idx]
    fc = o["force_cube"].double().cpu().numpy().T[idx]
    pot = o["pot"].double().cpu().numpy()[idx]
    de = abs(e - big["E"]) / abs(big["E"])
    er, ec, ep = orc.erms(big["fs"], f), orc.erms(big["fc"], fc), orc.erms(big["pot"], pot)
    print(f"1M FP32: dE/E {de:.2e} Erms(sphere) {er:.2e} Erms(cube) {ec:.2e} "
          f"Erms(phi_i) {ep:.2e}", flush=True)
    assert de <= 1e-5
    assert er <= 1e-5
    assert ec <= 1e-5
    assert ep <= 1e-5


def test_adjoint_fp32_vs_oracle(big):
    g, idx = big["g"], big["idx"]
    a = g.adjoint(big["x"], want_virial=False)
    torch.cuda.synchronize()
    er = orc.erms(-big["fc"], a["dEdr"].double().cpu().numpy().T[idx])
    eq = orc.erms(big["pot"], a["dEdq"].double().cpu().numpy()[idx])
    print(f"1M FP32 adjoint: Erms(dE/dr) {er:.2e} Erms(dE/dq) {eq:.2e}", flush=True)
    assert er <= 1e-5
    assert eq <= 1e-5


def test_rerun_and_permutation_bit_identical(big):
    g, x = big["g"], big["x"]
    o1 = g.step(x, force=True, potential=True)
    rho1 = g.getRhoS()
    f1 = o1["force"].clone()
    p1 = o1["pot"].clone()
    e1 = o1["energy"].clone()
    o2 = g.step(x, force=True, potential=True)
    assert torch.equal(g.getRhoS(), rho1)
    assert torch.equal(o2["force"], f1) and torch.equal(o2["energy"], e1)
    perm = torch.randperm(x.shape[0], generator=torch.Generator().manual_seed(3)).cuda()
    o3 = g.step(x[perm].contiguous(), force=True, potential=True)
    assert torch.equal(g.getRhoS(), rho1)
    assert torch.equal(o3["energy"], e1)
    assert torch.equal(o3["force"], f1[:, perm])
    assert torch.equal(o3["pot"], p1[perm])


def test_parseval_real_space_energy(big):
    g, x = big["g"], big["x"]
    ek = float(g.step(x, force=False, potential=False)["energy"].cpu()[0])
    g.spreadCharge1(x.shape[0], x)
    g.solvePoisson()
    er = g.calculateEnergy()
    print(f"1M FP32: k-space E {ek:.6f} real-space E {er:.6f}", flush=True)
    assert abs(ek - er) / abs(er) <= 1e-5


def test_water48k_fp64_vs_oracle():
    """The 48,000-charge water box (16,000 triplets, L = 78.3, M = 80) in FP64
    mode, the drop-in's own precision: energy, reference forces (sphere
    support, x ccelec) and per-atom potentials within 1e-10 of the FP64
    oracle's full step on the same particles."""
    xyzq, L, M, sigma = systems.make("water_48k_M80")
    p = orc.Params(M, L, sigma)
    print("oracle: FP64 step of 48,000 charges at M=80", flush=True)
    ref = orc.gse_step(p, xyzq, mode=0)
    g = api.CpuGSEk(torch.float64, torch.float64, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    o = g.step(torch.from_numpy(xyzq).cuda(), force=True, potential=True)
    torch.cuda.synchronize()
    e = float(o["energy"].cpu()[0])
    f = o["force"].cpu().numpy().T
    pot = o["pot"].cpu().numpy()
    de = abs(e - ref.energy) / abs(ref.energy)
    df = np.max(np.abs(f - ref.force)) / np.max(np.abs(ref.force))
    dp = np.max(np.abs(pot - ref.pot)) / np.max(np.abs(ref.pot))
    print(f"48k FP64: dE/E {de:.2e} max dF {df:.2e} max dphi_i {dp:.2e}", flush=True)
    assert de <= 1e-10
    assert df <= 1e-10
    assert dp <= 1e-10
