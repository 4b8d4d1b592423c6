"""The fused z pass of the FFT solve (2D x-y rocFFT transforms around one
kernel that runs the z transform, the G(k) multiply with the FP64
energy/virial sums and the inverse z transform) against the FP64 oracle and
against the 3D-transform form of the same solve.

Tolerances: FP64 grids/forces 1e-10 relative to max|oracle| (the Stockham z
transform differs from rocFFT's only by rounding, ~1e-15), energies 1e-10
relative; FP32: fused vs 3D form, force RMS (cpu_gse.cpp:131-146 Erms) 1e-5.
"""
import contextlib

import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

F64 = torch.float64
F32 = torch.float32
DEV = "cuda"
TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def to_dev(a, dt=F64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=DEV, dtype=dt).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.detach().double().cpu().numpy()


@contextlib.contextmanager
def fused(on, four=0, xy=1):
    old = api.get_tuning()
    api.set_tuning(fused_kz=int(on), zpass4=int(four), xy_plane=int(xy))
    try:
        yield
    finally:
        api.set_tuning(fused_kz=old["fused_kz"], zpass4=old["zpass4"], xy_plane=old["xy_plane"])


# Z factorisations: radix 8 only, 8*8*2, 6*4, 7*5, 3*3*3*... (27), 2*2 with
# odd X, and primes above 8 (no fused plan: the 3D form is kept)
DIMS = [
    ((16, 16, 64), (16.0, 16.5, 64.0), "fused"),
    ((12, 10, 128), (12.0, 10.5, 128.0), "fused"),
    ((18, 16, 24), (18.0, 16.2, 21.6), "fused"),
    ((9, 14, 35), (9.5, 14.0, 35.0), "fused"),
    ((20, 6, 27), (20.0, 6.5, 27.0), "fused"),
    ((15, 9, 4), (15.0, 9.5, 4.5), "fused"),
    ((10, 12, 11), (10.0, 12.0, 11.0), "3d"),
    ((8, 8, 13), (8.0, 8.0, 13.0), "3d"),
]


@pytest.mark.parametrize("xy", [1, 0], ids=["plane", "rocfft2d"])
@pytest.mark.parametrize("kind,sigma", [(0, 0.9487), (0, 0.0), (1, 0.9487), (2, 0.0)])
@pytest.mark.parametrize("dims,box,form", DIMS, ids=["x".join(map(str, d[0])) for d in DIMS])
def test_fused_solver_fp64_vs_oracle(kind, sigma, dims, box, form, xy):
    X, Y, Z = dims
    rho = np.random.default_rng(X * 1000 + Z).normal(size=(Z, Y, X))
    gf = [api.CpuGreensFuncGSE, api.CpuGreensFuncG2, api.CpuGreensFuncHermitian][kind]()
    if kind == 2:
        rho -= rho.mean()
        ref = orc.mehrstellen_solve(dims, box, rho)
    else:
        ref = orc.fft_solve(kind, dims, sigma, box, rho)
    out = {}
    for on in (True, False):
        with fused(on, 0, xy):
            s = api.CpuFFTSolver(gf, F64, *dims)
        want = form + ("/xy" if xy and form == "fused" else "")
        assert s.form() == (want if on else "3d")
        phi = torch.zeros((Z, Y, X), dtype=F64, device=DEV)
        s.run(phi, to_dev(rho), sigma, *box)
        out[on] = host(phi)
        if kind == 2:
            out[on] -= out[on].mean()
        assert rel(out[on], ref) <= TOL, (on, rel(out[on], ref))
    assert rel(out[True], out[False]) <= 1e-12


STEP_CASES = [
    ("water_64", lambda: systems.water_box(700, 27.0, seed=21), (28, 28, 64), (27.0, 27.0, 60.0)),
    ("water_48", lambda: systems.water_box(1000, 31.0, seed=3), (32, 32, 48), (31.0,) * 3),
    ("aniso_35", lambda: systems.water_box(150, 18.0, seed=7) * np.array([1.0, 0.9, 1.2, 1.0]),
     (18, 16, 35), (18.0, 16.2, 21.6)),
]


@pytest.mark.parametrize("name,make,dims,box", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_fused_gsek_step_fp64_vs_oracle(name, make, dims, box):
    xyzq = make()
    p = orc.Params(dims, box, 2.12)
    r = orc.gse_step(p, xyzq, mode=0)
    res = {}
    for on in (True, False):
        with fused(on):
            gs = api.CpuGSEk(F64, F64, *dims, p.sigma, 0.4, 3.0, 3.0, *box)
        assert gs.layout()["kspace"] == ("fused/xy" if on else "3d")
        o = gs.step(to_dev(xyzq), mode=0, force=True, potential=True, virial=True)
        E = float(host(o["energy"])[0])
        assert abs(E - r.energy) <= TOL * abs(r.energy), (on, E, r.energy)
        assert rel(host(o["force"]).T, r.force) <= TOL
        assert rel(host(o["pot"]), r.pot) <= TOL
        assert rel(host(o["virial"]).reshape(3, 3), r.virial) <= TOL
        res[on] = (E, host(o["force"]), host(o["virial"]))
        # the step repeated on the same handle gives the same result
        o2 = gs.step(to_dev(xyzq), mode=0, force=True, potential=True, virial=True)
        assert float(host(o2["energy"])[0]) == E
    assert abs(res[True][0] - res[False][0]) <= 1e-12 * abs(res[False][0])


def test_fused_gsek_fp32_matches_3d_form():
    xyzq = systems.water_box(4000, 49.0, seed=5)
    dims, box = (48, 48, 48), (49.0,) * 3
    out = {}
    for on in (True, False):
        with fused(on):
            gs = api.CpuGSEk(F32, F32, *dims, 2.12, 0.4, 3.0, 3.0, *box)
        o = gs.step(to_dev(xyzq, F32), mode=0, force=True, potential=False, virial=True)
        out[on] = (float(host(o["energy"])[0]), host(o["force"]), host(o["virial"]))
    fa, fb = out[True][1], out[False][1]
    erms = np.sqrt(np.sum((fa - fb) ** 2) / np.sum(fb ** 2))
    assert erms <= 1e-5, erms
    assert abs(out[True][0] - out[False][0]) <= 1e-5 * abs(out[False][0])
    assert rel(out[True][2], out[False][2]) <= 1e-5


def test_fused_gsek_mode_r_fp64():
    xyzq = systems.water_box(300, 24.0, seed=14)
    p = orc.Params((24, 24, 40), (24.0, 24.0, 40.0), 2.12)
    r = orc.gse_step(p, xyzq, mode=1)
    with fused(True):
        gs = api.CpuGSEk(F64, F64, *p.dims, p.sigma, 0.4, 3.0, 3.0, *p.box)
    assert gs.layout()["kspace"] == "fused/xy"
    o = gs.step(to_dev(xyzq), mode=1, force=True, potential=True)
    E = float(host(o["energy"])[0])
    assert abs(E - r.energy) <= TOL * abs(r.energy)
    assert rel(host(o["force"]).T, r.force) <= TOL
    assert rel(host(o["pot"]), r.pot) <= TOL


# The four-step z pass (registers, two LDS exchanges per direction) is built
# for Z = 80 (8 x 10) and Z = 216 (12 x 18); ragged column counts (Xh Y not a
# multiple of the 16 columns per workgroup) included.
FOUR_DIMS = [
    ((16, 18, 80), (16.0, 18.5, 80.0)),
    ((11, 7, 80), (11.0, 7.5, 80.0)),
    ((12, 10, 216), (12.0, 10.5, 216.0)),
]


@pytest.mark.parametrize("cols", [16, 32])
@pytest.mark.parametrize("xy", [1, 0], ids=["plane", "rocfft2d"])
@pytest.mark.parametrize("kind,sigma", [(0, 0.9487), (1, 0.9487), (2, 0.0)])
@pytest.mark.parametrize("dims,box", FOUR_DIMS, ids=["x".join(map(str, d[0])) for d in FOUR_DIMS])
@pytest.mark.parametrize("dt", [F64, F32], ids=["f64", "f32"])
def test_four_step_solver_vs_oracle_and_stockham(kind, sigma, dims, box, dt, xy, cols):
    X, Y, Z = dims
    rho = np.random.default_rng(X * 1000 + Z + 1).normal(size=(Z, Y, X))
    gf = [api.CpuGreensFuncGSE, api.CpuGreensFuncG2, api.CpuGreensFuncHermitian][kind]()
    if kind == 2:
        rho -= rho.mean()
        ref = orc.mehrstellen_solve(dims, box, rho)
    else:
        ref = orc.fft_solve(kind, dims, sigma, box, rho)
    tol = TOL if dt == F64 else 2e-6
    out = {}
    # X = 11 has no radix plan: the x-y transforms stay with rocFFT
    sfx = "/xy" if xy and X != 11 else ""
    old_cols = api.get_tuning()["zpass4_cols"]
    for four in (1, 0):
        api.set_tuning(zpass4_cols=cols)
        try:
            with fused(True, four, xy):
                s = api.CpuFFTSolver(gf, dt, *dims)
        finally:
            api.set_tuning(zpass4_cols=old_cols)
        assert s.form() == ("fused4" if four else "fused") + sfx
        phi = torch.zeros((Z, Y, X), dtype=dt, device=DEV)
        s.run(phi, to_dev(rho, dt), sigma, *box)
        out[four] = host(phi)
        if kind == 2:
            out[four] -= out[four].mean()
        assert rel(out[four], ref) <= tol, (four, rel(out[four], ref))
    assert rel(out[1], out[0]) <= (1e-12 if dt == F64 else 2e-6)


def test_four_step_gsek_step_fp64_vs_oracle():
    xyzq = systems.water_box(400, 24.0, seed=31)
    dims, box = (24, 24, 80), (24.0, 24.0, 72.0)
    p = orc.Params(dims, box, 2.12)
    for mode in (0, 1):
        r = orc.gse_step(p, xyzq, mode=mode)
        with fused(True, 1):
            gs = api.CpuGSEk(F64, F64, *dims, p.sigma, 0.4, 3.0, 3.0, *box)
        assert gs.layout()["kspace"] == "fused4/xy"
        o = gs.step(to_dev(xyzq), mode=mode, force=True, potential=True, virial=mode == 0)
        E = float(host(o["energy"])[0])
        assert abs(E - r.energy) <= TOL * abs(r.energy), (mode, E, r.energy)
        assert rel(host(o["force"]).T, r.force) <= TOL
        assert rel(host(o["pot"]), r.pot) <= TOL
        if mode == 0:
            assert rel(host(o["virial"]).reshape(3, 3), r.virial) <= TOL


def test_four_step_gsek_fp32_water_48k_M80():
    """The FP32 step at 48k charges on the 80^3 grid (the default four-step
    z pass): force Erms (cpu_gse.cpp:131-146) against the FP64 oracle
    <= 1e-5, energy and virial against the Stockham z pass to 1e-5."""
    xyzq, L, M, sigma = systems.make("water_48k_M80")
    p = orc.Params(M, L, sigma)
    ref = orc.gse_step(p, xyzq.astype(np.float32).astype(np.float64))
    out = {}
    for four in (1, 0):
        with fused(True, four, 0):
            gs = api.CpuGSEk(F32, F32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
        assert gs.layout()["kspace"] == ("fused4" if four else "fused") + "/y4"
        o = gs.step(to_dev(xyzq, F32), force=True, potential=True, virial=True)
        out[four] = (float(host(o["energy"])[0]), host(o["force"]), host(o["virial"]))
        erms = orc.erms(ref.force, out[four][1].T)
        assert erms <= 1e-5, (four, erms)
    assert abs(out[1][0] - out[0][0]) <= 1e-5 * abs(out[0][0])
    assert rel(out[1][2], out[0][2]) <= 1e-5


# Plane kernels (one workgroup per z plane does both x-y transforms in LDS)
# on edge shapes: lengths 2 and 3, radix 7 twice (49), odd X and Y, and a
# plane that fits the 160 KiB LDS budget in FP32 but not in FP64 (there the
# x-y transforms stay with rocFFT).
PLANE_DIMS = [
    ((2, 2, 4), (2.0, 2.0, 4.0)),
    ((3, 5, 8), (3.0, 5.0, 8.0)),
    ((7, 49, 6), (7.0, 49.0, 6.0)),
    ((21, 10, 12), (21.0, 10.5, 12.0)),
    ((128, 96, 16), (128.0, 96.0, 16.0)),
]


@pytest.mark.parametrize("dims,box", PLANE_DIMS, ids=["x".join(map(str, d[0])) for d in PLANE_DIMS])
@pytest.mark.parametrize("dt", [F64, F32], ids=["f64", "f32"])
def test_plane_kernels_vs_oracle_and_rocfft2d(dims, box, dt):
    X, Y, Z = dims
    rho = np.random.default_rng(X * 7 + Y * 3 + Z).normal(size=(Z, Y, X))
    ref = orc.fft_solve(0, dims, 0.9487, box, rho)
    tol = TOL if dt == F64 else 2e-6
    plane_fits = not (dt == F64 and dims[0] == 128)
    out = {}
    for xy in (1, 0):
        with fused(True, 0, xy):
            s = api.CpuFFTSolver(api.CpuGreensFuncGSE(), dt, *dims)
        assert s.form() == ("fused/xy" if xy and plane_fits else "fused")
        phi = torch.zeros((Z, Y, X), dtype=dt, device=DEV)
        s.run(phi, to_dev(rho, dt), 0.9487, *box)
        out[xy] = host(phi)
        assert rel(out[xy], ref) <= tol, (xy, rel(out[xy], ref))
    assert rel(out[1], out[0]) <= (1e-12 if dt == F64 else 2e-6)


# Four-step y pass (Y = 216 = 12 x 18 in registers, 32 columns per
# workgroup; Y = 80 = 8 x 10, 32 columns in FP32, 16 in FP64) between 1D x rocFFT
# transforms, around both z passes; odd X and column counts Xh Z that are
# not a multiple of the columns per workgroup included.
YPASS_DIMS = [
    ((16, 80, 24), (16.0, 80.0, 24.5)),
    ((9, 80, 80), (9.0, 80.0, 80.0)),
    ((12, 216, 10), (12.0, 216.0, 10.5)),
    ((7, 216, 80), (7.0, 216.0, 80.0)),
    ((20, 216, 216), (20.0, 216.0, 216.0)),
]


@pytest.mark.parametrize("kind,sigma", [(0, 0.9487), (2, 0.0)])
@pytest.mark.parametrize("dims,box", YPASS_DIMS, ids=["x".join(map(str, d[0])) for d in YPASS_DIMS])
@pytest.mark.parametrize("dt", [F64, F32], ids=["f64", "f32"])
def test_ypass_solver_vs_oracle_and_rocfft2d(kind, sigma, dims, box, dt):
    X, Y, Z = dims
    rho = np.random.default_rng(X * 5 + Z).normal(size=(Z, Y, X))
    gf = [api.CpuGreensFuncGSE, api.CpuGreensFuncG2, api.CpuGreensFuncHermitian][kind]()
    if kind == 2:
        rho -= rho.mean()
        ref = orc.mehrstellen_solve(dims, box, rho)
    else:
        ref = orc.fft_solve(kind, dims, sigma, box, rho)
    tol = TOL if dt == F64 else 2e-6
    four = Z in (80, 216)
    built = Y in (80, 216)
    old = api.get_tuning()["ypass4"]
    out = {}
    for yp in (1, 0):
        api.set_tuning(ypass4=yp)
        try:
            with fused(True, 1, 0):
                s = api.CpuFFTSolver(gf, dt, *dims)
        finally:
            api.set_tuning(ypass4=old)
        assert s.form() == ("fused4" if four else "fused") + ("/y4" if yp and built else "")
        phi = torch.zeros((Z, Y, X), dtype=dt, device=DEV)
        s.run(phi, to_dev(rho, dt), sigma, *box)
        out[yp] = host(phi)
        if kind == 2:
            out[yp] -= out[yp].mean()
        assert rel(out[yp], ref) <= tol, (yp, rel(out[yp], ref))
    assert rel(out[1], out[0]) <= (1e-12 if dt == F64 else 2e-6)
