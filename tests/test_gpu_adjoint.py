"""Adjoint sweep (dE/dq, dE/dr, virial) behind torch.autograd, on the 3,000-
charge water box (L = 31.0, M = 32, sigma = 2.12).

Checks:
  * autograd gradients vs the FP64 oracle: dE/dr = -(cube force), dE/dq =
    phi_i, to 1e-10 relative;
  * central finite differences of the GPU FP64 energy (h = 1e-5 A for
    positions, 1e-5 e for charges, box strain 1e-6 for the diagonal virial);
  * the reference force path: -dE/dr minus the reference sphere force equals
    the oracle's corner term (cube minus sphere) to 1e-10 of max|F|;
  * upstream-gradient scaling and torch.autograd.gradcheck on a small box;
  * FP32 adjoint vs the FP64 oracle: Erms <= 1e-5.
"""
import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc
from tests.autograd_gse import energy

pytestmark = pytest.mark.gpu

F64, F32 = torch.float64, torch.float32


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def water3k():
    xyzq, L, M, sigma = systems.make("water_3k_M32")
    p = orc.Params(M, L, sigma)
    ref = orc.gse_step(p, xyzq)
    return xyzq, L, M, sigma, ref


def gsek(dt, L, M, sigma):
    return api.CpuGSEk(dt, dt, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)


def test_autograd_matches_oracle_fp64(water3k):
    xyzq, L, M, sigma, ref = water3k
    g = gsek(F64, L, M, sigma)
    x = torch.from_numpy(xyzq).cuda().requires_grad_(True)
    e = energy(x, g)
    e.backward()
    grad = x.grad.cpu().numpy()
    print(f"E gpu {float(e):.12f} oracle {ref.energy:.12f}", flush=True)
    assert abs(float(e) - ref.energy) / abs(ref.energy) < 1e-10
    assert rel(grad[:, :3], -ref.force_cube) < 1e-10
    assert rel(grad[:, 3], ref.pot) < 1e-10


def test_adjoint_virial_matches_oracle_fp64(water3k):
    xyzq, L, M, sigma, ref = water3k
    g = gsek(F64, L, M, sigma)
    a = g.adjoint(torch.from_numpy(xyzq).cuda())
    torch.cuda.synchronize()
    assert rel(a["virial"].cpu().numpy().reshape(3, 3), ref.virial) < 1e-10


def test_finite_differences_fp64(water3k):
    xyzq, L, M, sigma, _ = water3k
    g = gsek(F64, L, M, sigma)
    x = torch.from_numpy(xyzq).cuda()

    def E(xx):
        return float(g.step(xx, force=False, potential=False)["energy"].cpu()[0])

    a = g.adjoint(x)
    dEdr = a["dEdr"].cpu().numpy()
    dEdq = a["dEdq"].cpu().numpy()
    scale_r = np.max(np.abs(dEdr))
    scale_q = np.max(np.abs(dEdq))
    rng = np.random.default_rng(11)
    h = 1e-5
    worst_r = worst_q = 0.0
    for i in rng.choice(len(xyzq), 8, replace=False):
        for d in range(4):
            xp, xm = x.clone(), x.clone()
            xp[i, d] += h
            xm[i, d] -= h
            fd = (E(xp) - E(xm)) / (2 * h)
            if d < 3:
                worst_r = max(worst_r, abs(fd - dEdr[d, i]) / scale_r)
            else:
                worst_q = max(worst_q, abs(fd - dEdq[i]) / scale_q)
    print(f"FD worst rel: dE/dr {worst_r:.2e} dE/dq {worst_q:.2e}", flush=True)
    assert worst_r < 1e-6
    assert worst_q < 1e-6
    # diagonal virial: W_aa = -dE/d(eps_aa), box and positions co-scaled
    W = a["virial"].cpu().numpy().reshape(3, 3)
    eps = 1e-6
    for d in range(3):
        vals = []
        for s in (1 + eps, 1 - eps):
            box = [L, L, L]
            box[d] *= s
            g.setBoxSize(*box)
            xs = x.clone()
            xs[:, d] *= s
            vals.append(E(xs))
        g.setBoxSize(L, L, L)
        fd = -(vals[0] - vals[1]) / (2 * eps)
        print(f"W[{d}{d}] adjoint {W[d, d]:.9f} FD {fd:.9f}", flush=True)
        assert abs(fd - W[d, d]) / np.max(np.abs(W)) < 1e-6


def test_reference_force_path(water3k):
    xyzq, L, M, sigma, ref = water3k
    g = gsek(F64, L, M, sigma)
    x = torch.from_numpy(xyzq).cuda()
    a = g.adjoint(x)
    o = g.step(x, force=True, potential=False)
    torch.cuda.synchronize()
    exact = -a["dEdr"].cpu().numpy().T
    sphere = o["force"].cpu().numpy().T
    corner_ref = ref.force_cube - ref.force
    fmax = np.max(np.abs(ref.force))
    assert np.max(np.abs((exact - sphere) - corner_ref)) / fmax < 1e-10
    # the reference force approximates the exact gradient (truncation-level)
    print(f"reference force vs exact gradient, Erms {orc.erms(exact, sphere):.3e}", flush=True)
    assert orc.erms(exact, sphere) < 1e-2


def test_upstream_gradient_scaling(water3k):
    xyzq, L, M, sigma, _ = water3k
    g = gsek(F64, L, M, sigma)
    x1 = torch.from_numpy(xyzq).cuda().requires_grad_(True)
    energy(x1, g).backward()
    x2 = torch.from_numpy(xyzq).cuda().requires_grad_(True)
    (2.5 * energy(x2, g)).backward()
    assert rel(x2.grad.cpu().numpy(), 2.5 * x1.grad.cpu().numpy()) < 1e-14


def test_gradcheck_small_box():
    xyzq = systems.random_alternating(10, 12.0, seed=21)
    g = api.CpuGSEk(F64, F64, 12, 12, 12, 2.0, 0.4, 3.0, 3.0, 12.0, 12.0, 12.0)
    x = torch.from_numpy(xyzq).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: energy(t, g), (x,), eps=1e-6, atol=1e-6,
                                    rtol=1e-5)


def test_adjoint_fp32_vs_fp64_oracle(water3k):
    xyzq, L, M, sigma, _ = water3k
    x32 = torch.from_numpy(xyzq).to("cuda", F32)
    # the oracle sees the same float32-rounded particles
    ref = orc.gse_step(orc.Params(M, L, sigma), x32.double().cpu().numpy())
    g = gsek(F32, L, M, sigma)
    a = g.adjoint(x32)
    torch.cuda.synchronize()
    er = orc.erms(-ref.force_cube, a["dEdr"].double().cpu().numpy().T)
    eq = orc.erms(ref.pot, a["dEdq"].double().cpu().numpy())
    ev = rel(a["virial"].cpu().numpy().reshape(3, 3), ref.virial)
    print(f"FP32 adjoint: Erms(dE/dr) {er:.2e} Erms(dE/dq) {eq:.2e} virial {ev:.2e}", flush=True)
    assert er <= 1e-5
    assert eq <= 1e-5
    assert ev <= 1e-5
