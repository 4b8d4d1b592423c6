"""The FP32 spread on the matrix cores (int8 limb products summed exactly per
bin, tuning knob spread_mma, off by default) against the FP64 oracle and
against the LDS-atomic spread.

The two FP32 kernels round differently (the matrix-core brick quantises the
per-axis factors to 23 bits and rounds each brick point once; the LDS path
rounds every contribution at the brick's scale), so they agree to FP32
accuracy, not bit for bit: grids within 2e-6 of max|grid| of each other and
of the oracle.  Each kernel on its own is deterministic: its grid is
bit-identical under any permutation of the particles and whichever particles
take the overflow path, because every sum it forms is an exact integer sum.
"""
import contextlib

import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

F32 = torch.float32
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def to_dev(a, dt=F32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=DEV, dtype=dt).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.detach().double().cpu().numpy()


@contextlib.contextmanager
def mma(on):
    old = api.get_tuning()["spread_mma"]
    api.set_tuning(spread_mma=int(on))
    try:
        yield
    finally:
        api.set_tuning(spread_mma=old)


def spread32(xyzq, p, on):
    with mma(on):
        g = api.CpuGaussCharge(F32, F32, *p.dims)
        rho = torch.zeros((p.Z, p.Y, p.X), dtype=F32, device=DEV)
        g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq), *p.box, rho)
        return rho, g.layout()


CASES = [
    ("water_3k_M32", lambda: systems.water_box(1000, 31.0, seed=3), 32, 31.0),
    ("water_48k_M80", lambda: systems.make("water_48k_M80")[0], 80, 78.3),
    ("water_10k_M48", lambda: systems.water_box(3400, 46.0, seed=9), 48, 46.0),
]


@pytest.mark.parametrize("name,make,M,L", CASES, ids=[c[0] for c in CASES])
def test_mma_spread_vs_oracle_and_lds_path(name, make, M, L):
    xyzq = make()
    p = orc.Params(M, L, 2.12)
    r1, lay1 = spread32(xyzq, p, True)
    r0, lay0 = spread32(xyzq, p, False)
    assert lay1["spread"] == "mma" and lay0["spread"] == "lds_atomic", (lay1, lay0)
    ref = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq.astype(np.float32).astype(np.float64), p.box)
    e1, e0 = rel(host(r1), ref), rel(host(r0), ref)
    print(f"{name}: rel err mma {e1:.3e} lds {e0:.3e}", flush=True)
    assert e1 <= 2e-6 and e0 <= 2e-6
    assert rel(host(r1), host(r0)) <= 2e-6
    # total charge is conserved to FP32 rounding
    assert abs(float(host(r1).sum()) - float(host(r0).sum())) <= 1e-5 * float(np.abs(host(r0)).sum())


def test_mma_spread_order_independent():
    xyzq = systems.water_box(1000, 31.0, seed=3)
    p = orc.Params(32, 31.0, 2.12)
    a, _ = spread32(xyzq, p, True)
    for seed in range(3):
        perm = np.random.default_rng(seed).permutation(len(xyzq))
        b, _ = spread32(xyzq[perm], p, True)
        assert torch.equal(a, b)


def test_mma_spread_overflow_cluster():
    """A cluster overfills its bin: the bin's overflow particles are summed
    into its own brick, so the grid does not depend on which of them
    overflowed (permutations change that) and stays close to the oracle."""
    rng = np.random.default_rng(17)
    n = 2000
    xyzq = np.empty((n, 4))
    xyzq[:, :3] = 10.0 + rng.random((n, 3)) * 0.9
    xyzq[:, 3] = 0.25 + 0.5 * rng.random(n)
    p = orc.Params(32, 31.0, 2.12)
    r1, lay = spread32(xyzq, p, True)
    assert lay["spread"] == "mma" and n > lay["cap_bin"]
    for seed in range(3):
        perm = np.random.default_rng(seed).permutation(n)
        b, _ = spread32(xyzq[perm], p, True)
        assert torch.equal(r1, b)
    ref = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq.astype(np.float32).astype(np.float64), p.box)
    assert rel(host(r1), ref) <= 1e-6


def test_mma_spread_giant_bin_takes_the_per_contribution_path():
    """A bin with more than 32768 particles is spread per contribution
    (int32 class sums would no longer be exact): with every particle in that
    bin, the grid is bit-identical to the LDS-atomic kernel's, which handles
    the same bin's overflow the same way."""
    rng = np.random.default_rng(5)
    n = 33000
    xyzq = np.empty((n, 4))
    xyzq[:, :3] = 10.0 + rng.random((n, 3)) * 0.9
    xyzq[:, 3] = np.where(np.arange(n) % 2 == 0, 0.4, -0.4)
    p = orc.Params(32, 31.0, 2.12)
    r1, lay1 = spread32(xyzq, p, True)
    r0, _ = spread32(xyzq, p, False)
    assert lay1["spread"] == "mma"
    assert torch.equal(r1, r0)


def test_mma_gsek_step_fp32_force_rms():
    """The FP32 step at 48k charges, M = 80 with the matrix-core spread:
    force Erms (cpu_gse.cpp:131-146) against the FP64 oracle <= 1e-5."""
    xyzq, L, M, sigma = systems.make("water_48k_M80")
    p = orc.Params(M, L, sigma)
    ref = orc.gse_step(p, xyzq.astype(np.float32).astype(np.float64))
    with mma(True):
        gs = api.CpuGSEk(F32, F32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
        o = gs.step(to_dev(xyzq), force=True, potential=True)
    assert gs.layout()["spread"] == "mma"
    erms = orc.erms(ref.force, host(o["force"]).T)
    de = abs(float(host(o["energy"])[0]) - ref.energy) / abs(ref.energy)
    print(f"mma step: Erms {erms:.3e} dE/E {de:.3e}", flush=True)
    assert erms <= 1e-5 and de <= 1e-5
