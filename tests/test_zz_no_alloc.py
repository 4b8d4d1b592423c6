"""The allocation-free spread (gse_gauss_charge_spread_charge_to_grid_noalloc)
against the allocating entry point it varies, through the C-ABI.

Both run the same binning, spread and conversion kernels, so the grids must
be bit-identical.  Two calls in a row on the shared scratch, the second with
a smaller system and grid, check that what the first call left behind does
not leak into the second; a request beyond the scratch limit is refused on
the host with a clear error.
"""
import ctypes

import numpy as np
import pytest
import torch

from gse_amd import _lib, api, systems

pytestmark = pytest.mark.gpu

F32, F64 = torch.float32, torch.float64


def _spread_both(dt, xyzq, M, L, sigma1=2.12 * 0.4 ** 0.5):
    rc = 3.0 * 2.0 ** 0.5 * sigma1
    x = torch.from_numpy(xyzq).to("cuda", dt).contiguous()
    g = api.CpuGaussCharge(dt, dt, M, M, M)
    ref = torch.zeros((M, M, M), dtype=dt, device="cuda")
    g.spreadChargeToGrid(sigma1, rc, len(xyzq), x, L, L, L, ref)
    out = torch.full((M, M, M), float("nan"), dtype=dt, device="cuda")
    g._s()
    L_ = _lib.load()
    st = L_.gse_gauss_charge_spread_charge_to_grid_noalloc(
        g._h, sigma1, rc, len(xyzq), ctypes.c_void_p(x.data_ptr()), L, L, L,
        ctypes.c_void_p(out.data_ptr()), 0)
    _lib.check(st)
    torch.cuda.synchronize()
    return ref, out


def test_noalloc_spread_matches_allocating_spread_twice():
    print("noalloc: 48,000 FP32 charges, M=80", flush=True)
    xa, La, Ma, _ = systems.make("water_48k_M80")
    ref_a, out_a = _spread_both(F32, xa, Ma, La)
    assert torch.equal(out_a, ref_a)
    print("noalloc: 3,000 FP32 charges, M=32 (second, smaller call)", flush=True)
    xb, Lb, Mb, _ = systems.make("water_3k_M32")
    ref_b, out_b = _spread_both(F32, xb, Mb, Lb)
    assert torch.equal(out_b, ref_b)
    print("noalloc: 3,000 FP64 charges, M=32", flush=True)
    ref_c, out_c = _spread_both(F64, xb, Mb, Lb)
    assert torch.equal(out_c, ref_c)


def test_noalloc_spread_with_overflow_cluster():
    """A dense cluster overfills its bin: the overflow list goes through the
    scratch's int64 grid, which must start at zero on every call."""
    rng = np.random.default_rng(5)
    x = systems.water_box(2000, 46.0, seed=5)
    cl = np.empty((500, 4))
    cl[:, :3] = 21.0 + rng.random((500, 3)) * 0.9
    cl[:, 3] = 0.1 * rng.standard_normal(500)
    x = np.concatenate([x, cl])
    for dt in (F32, F64, F32):
        ref, out = _spread_both(dt, x, 48, 46.0)
        assert torch.equal(out, ref)


def test_noalloc_spread_rejects_requests_beyond_the_scratch():
    """FP64 bricks of the 48k M=80 box need more than the 64 MiB limit: the
    call fails on the host before any device work and names the limit."""
    xa, La, Ma, _ = systems.make("water_48k_M80")
    x = torch.from_numpy(xa).to("cuda", F64).contiguous()
    g = api.CpuGaussCharge(F64, F64, Ma, Ma, Ma)
    rho = torch.zeros((Ma, Ma, Ma), dtype=F64, device="cuda")
    L_ = _lib.load()
    s1 = 2.12 * 0.4 ** 0.5
    st = L_.gse_gauss_charge_spread_charge_to_grid_noalloc(
        g._h, s1, 3.0 * 2.0 ** 0.5 * s1, len(xa), ctypes.c_void_p(x.data_ptr()), La, La, La,
        ctypes.c_void_p(rho.data_ptr()), 0)
    assert st == 1
    assert b"limit" in L_.gse_last_error()
    # the handle and the shared scratch still work afterwards
    xb, Lb, Mb, _ = systems.make("water_3k_M32")
    ref, out = _spread_both(F64, xb, Mb, Lb)
    assert torch.equal(out, ref)
