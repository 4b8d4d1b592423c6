"""HIP path vs the FP64 oracle, through the C-ABI (gse_amd/api.py -> ctypes).

Tolerances (the bar for FP64 mode is 1e-10 relative):
  * grids, forces, potentials: max|gpu - oracle| / max|oracle| <= 1e-10
    (the spread's int64 fixed point and the FFT differ from the oracle only
    by rounding, ~1e-14);
  * energies: |gpu - oracle| / |oracle| <= 1e-10;
  * FP32 mode: relative force RMS (cpu_gse.cpp:131-146 Erms) <= 1e-5.
"""
import math

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
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b)))


def to_dev(a, dt=F64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=DEV, dtype=dt).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.detach().double().cpu().numpy()


def grid(dims, dt=F64):
    X, Y, Z = dims
    return torch.zeros((Z, Y, X), dtype=dt, device=DEV)


CASES = [
    # name, xyzq factory, dims, box, sigma(total)
    ("pair_diag", lambda: systems.pair(1.0, 20.0, orientation="diagonal"), (20, 20, 20), (20.0,) * 3, 2.121320344),
    ("pair_seed1", lambda: systems.pair(1.0, 20.0, seed=1), (20, 20, 20), (20.0,) * 3, 2.121320344),
    ("water_3k", lambda: systems.water_box(1000, 31.0, seed=3), (32, 32, 32), (31.0,) * 3, 2.12),
    ("aniso", lambda: systems.water_box(150, 18.0, seed=7) * np.array([1.0, 0.9, 1.2, 1.0]), (18, 16, 24), (18.0, 16.2, 21.6), 2.0),
    ("wrap_M8", lambda: systems.random_alternating(12, 8.0, seed=4), (8, 8, 8), (8.0,) * 3, 2.121320344),
    ("negative", lambda: systems.random_alternating(40, 16.0, seed=6) - np.array([8.0, 3.0, 15.0, 0.0]), (16, 16, 16), (16.0,) * 3, 2.0),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, make, dims, box, sigma = request.param
    xyzq = make()
    p = orc.Params(dims, box, sigma)
    return name, xyzq, p


# --------------------------------------------------------------- spread
@pytest.mark.parametrize("normalize", [False, True])
def test_spread_fp64(case, normalize):
    name, xyzq, p = case
    g = api.CpuGaussCharge(F64, F64, *p.dims)
    rho = grid(p.dims)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq), *p.box, rho, normalize)
    ref = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq, p.box, normalize)
    assert rel(host(rho), ref) <= TOL, name
    # bit-identical on a second run (int64 accumulation)
    rho2 = grid(p.dims)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq), *p.box, rho2, normalize)
    assert torch.equal(rho, rho2)


def test_spread_order_independent():
    """Reversing the particle order leaves the grid bit-identical."""
    xyzq = systems.water_box(500, 25.0, seed=12)
    p = orc.Params(24, 25.0, 2.12)
    g = api.CpuGaussCharge(F64, F64, *p.dims)
    a, b = grid(p.dims), grid(p.dims)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq), *p.box, a)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq[::-1]), *p.box, b)
    assert torch.equal(a, b)


def test_spread_fp32_vs_oracle():
    xyzq = systems.water_box(1000, 31.0, seed=3)
    p = orc.Params(32, 31.0, 2.12)
    g = api.CpuGaussCharge(F32, F32, *p.dims)
    rho = grid(p.dims, F32)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq, F32), *p.box, rho)
    ref = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq.astype(np.float32).astype(np.float64), p.box)
    assert rel(host(rho), ref) < 2e-6


# --------------------------------------------------------------- gather
def test_interpolate_force_fp64(case):
    name, xyzq, p = case
    X, Y, Z = p.dims
    phi = np.random.default_rng(5).normal(size=(Z, Y, X))
    g = api.CpuGaussCharge(F64, F64, *p.dims)
    n = len(xyzq)
    f = torch.zeros((3, n), dtype=F64, device=DEV)
    g.interpolateForce(p.sigma1, p.rcut1, n, to_dev(xyzq), *p.box, to_dev(phi), f[0], f[1], f[2])
    ref = orc.interpolate(p.dims, p.sigma1, p.rcut1, xyzq, p.box, phi)
    assert rel(host(f).T, ref) <= TOL, name


def test_gather_cube_potential_virial_fp64(case):
    name, xyzq, p = case
    X, Y, Z = p.dims
    phi = np.random.default_rng(6).normal(size=(Z, Y, X))
    g = api.CpuGaussCharge(F64, F64, *p.dims)
    o = g.gather(p.sigma1, p.rcut1, to_dev(xyzq), p.box, to_dev(phi), 15)
    pot, fc, vir = orc.gather_cube(p.dims, p.sigma1, p.rcut1, xyzq, p.box, phi)
    fs = orc.interpolate(p.dims, p.sigma1, p.rcut1, xyzq, p.box, phi)
    assert rel(host(o["force"]).T, fs) <= TOL
    assert rel(host(o["force_cube"]).T, fc) <= TOL
    assert rel(host(o["pot"]), pot) <= TOL
    assert rel(host(o["virial"]).reshape(3, 3), vir) <= TOL


# ---------------------------------------------------------- k-space
@pytest.mark.parametrize("kind,sigma", [(0, 0.9487), (0, 0.0), (1, 0.9487)])
@pytest.mark.parametrize("dims,box", [((20, 20, 20), (20.0,) * 3), ((18, 16, 24), (18.0, 16.2, 21.6)),
                                      ((15, 9, 10), (15.0, 9.5, 10.5))])
def test_fft_solver_fp64(kind, sigma, dims, box):
    X, Y, Z = dims
    rho = np.random.default_rng(8).normal(size=(Z, Y, X))
    gf = api.CpuGreensFuncGSE() if kind == 0 else api.CpuGreensFuncG2()
    s = api.CpuFFTSolver(gf, F64, *dims)
    phi = grid(dims)
    s.run(phi, to_dev(rho), sigma, *box)
    ref = orc.fft_solve(kind, dims, sigma, box, rho)
    assert rel(host(phi), ref) <= TOL


def test_greens_mul_and_transforms():
    dims, box = (12, 10, 9), (12.0, 10.5, 9.5)
    X, Y, Z = dims
    rho = np.random.default_rng(9).normal(size=(Z, Y, X))
    s = api.CpuFFTSolver(api.CpuGreensFuncGSE(), F64, *dims)
    spec = torch.zeros((Z, Y, X // 2 + 1), dtype=torch.complex128, device=DEV)
    s.r2c(to_dev(rho), spec)
    ref_spec = orc.fft_r2c(rho)
    assert rel(host(torch.view_as_real(spec)), np.stack([ref_spec.real, ref_spec.imag], -1)) <= TOL
    api.CpuGreensFuncGSE().mul(X, Y, Z, *box, 0.9, spec)
    ref_mul = orc.greens_mul(0, dims, box, 0.9, ref_spec)
    assert rel(host(torch.view_as_real(spec)), np.stack([ref_mul.real, ref_mul.imag], -1)) <= TOL
    phi = grid(dims)
    s.c2r(spec, phi)
    assert rel(host(phi), orc.fft_c2r(ref_mul, X)) <= TOL


# ------------------------------------------------------- convolution
def test_spread_on_grid_and_convolution_fp64():
    dims, box, sigma = (16, 12, 20), (16.5, 12.0, 19.0), 0.9487
    X, Y, Z = dims
    rcut = 3 * math.sqrt(2) * sigma
    rho = np.random.default_rng(10).normal(size=(Z, Y, X))
    g = api.CpuGaussCharge(F64, F64, *dims)
    out = grid(dims)
    g.spreadChargeOnGrid(sigma, rcut, *box, to_dev(rho), out)
    assert rel(host(out), orc.spread_on_grid(dims, sigma, rcut, box, rho)) <= TOL
    R = 5
    k = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * 2.12 ** 2))
    c = api.CpuConvolution(F64, *dims)
    c.setupKernel(R, R, R, k, k * 0.5, k * 2.0)
    o1 = grid(dims)
    c.run(to_dev(rho), o1)
    assert rel(host(o1), orc.conv_run_x(dims, k, rho)) <= TOL
    o3 = grid(dims)
    c.runXYZ(to_dev(rho), o3)
    assert rel(host(o3), orc.conv_run_xyz(dims, k, k * 0.5, k * 2.0, rho)) <= TOL


def test_convolution_fp32_compensated():
    dims = (32, 32, 32)
    X, Y, Z = dims
    rho = np.random.default_rng(11).normal(size=(Z, Y, X))
    R = 5
    k = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * 2.12 ** 2))
    c = api.CpuConvolution(F32, *dims)
    c.setupKernel(R, R, R, k, k, k)
    o = grid(dims, F32)
    c.runXYZ(to_dev(rho, F32), o)
    ref = orc.conv_run_xyz(dims, k, k, k, rho.astype(np.float32).astype(np.float64))
    assert rel(host(o), ref) < 1e-6


def test_efield_dipole_fp64():
    dims, box = (10, 12, 8), (10.0, 12.5, 8.5)
    X, Y, Z = dims
    phi = np.random.default_rng(12).normal(size=(Z, Y, X))
    g = api.CpuGaussCharge(F64, F64, *dims)
    e = [grid(dims) for _ in range(3)]
    g.calcElectricFieldOnGrid(*box, to_dev(phi), *e)
    ref = orc.efield(dims, box, phi)
    for a in range(3):
        assert rel(host(e[a]), ref[a]) <= TOL
    d = g.calcDipoleSum(*box, to_dev(phi))
    assert rel(d, orc.dipole_sum(dims, box, phi)) <= TOL


def test_support_size_matches_oracle():
    g = api.CpuGaussCharge(F64, F64, 20, 20, 20)
    assert g.calcSupportSize(5.6921, 20.0, 20.0, 20.0) == orc.support_size((20, 20, 20), 5.6921, (20.0,) * 3)


# ------------------------------------------------------ full GSE step
def test_gsek_step_fp64(case):
    name, xyzq, p = case
    r = orc.gse_step(p, xyzq, mode=0)
    gs = api.CpuGSEk(F64, F64, *p.dims, p.sigma, 0.4, 3.0, 3.0, *p.box)
    o = gs.step(to_dev(xyzq), mode=0, force=True, potential=True, cube=True, virial=True)
    E = float(host(o["energy"])[0])
    assert abs(E - r.energy) <= TOL * abs(r.energy), (name, E, r.energy)
    assert rel(host(o["force"]).T, r.force) <= TOL
    assert rel(host(o["force_cube"]).T, r.force_cube) <= TOL
    assert rel(host(o["pot"]), r.pot) <= TOL
    assert rel(host(o["virial"]).reshape(3, 3), r.virial) <= TOL


def test_gsek_reference_call_sequence_fp64():
    """The CpuGSEk calls in the order cpu_gse.cpp:719-735 makes them."""
    xyzq = systems.pair(1.0, 20.0, orientation="diagonal")
    p = orc.Params(20, 20.0, 2.121320344)
    r = orc.gse_step(p, xyzq, mode=0, want_grids=True)
    gs = api.CpuGSEk(F64, F64, 20, 20, 20, p.sigma, 0.4, 3.0, 3.0, 20.0, 20.0, 20.0)
    assert abs(gs.getSigma1() - p.sigma1) < 1e-15 and abs(gs.getSigma2() - p.sigma2) < 1e-15
    x = to_dev(xyzq)
    gs.spreadCharge1(2, x)
    gs.solvePoisson()
    E = gs.calculateEnergy()
    f = torch.zeros((3, 2), dtype=F64, device=DEV)
    gs.interpolateForce(2, x, f[0], f[1], f[2])
    assert abs(E - r.energy) <= TOL * abs(r.energy)
    assert rel(host(f).T, r.force) <= TOL
    assert rel(host(gs.getRhoS()), r.rho) <= TOL
    assert rel(host(gs.getPhi()), r.phi) <= TOL


def test_gsek_mode_r_fp64():
    xyzq = systems.water_box(300, 24.0, seed=14)
    p = orc.Params(24, 24.0, 2.12)
    r = orc.gse_step(p, xyzq, mode=1)
    gs = api.CpuGSEk(F64, F64, *p.dims, p.sigma, 0.4, 3.0, 3.0, *p.box)
    o = gs.step(to_dev(xyzq), mode=1, force=True, potential=True, cube=True)
    E = float(host(o["energy"])[0])
    assert abs(E - r.energy) <= TOL * abs(r.energy)
    assert rel(host(o["force"]).T, r.force) <= TOL
    assert rel(host(o["pot"]), r.pot) <= TOL


def test_gsek_fp32_water48k_force_rms():
    """FP32 spread/FFT/gather with FP64 G table and reductions, 48,000
    charges at M=80, against the FP64 oracle on the same inputs."""
    xyzq, L, M, sigma = systems.make("water_48k_M80")
    p = orc.Params(M, L, sigma)
    r = orc.gse_step(p, xyzq, mode=0)
    gs = api.CpuGSEk(F32, F32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    o = gs.step(to_dev(xyzq, F32), mode=0, force=True, potential=True, cube=True)
    erms = orc.erms(r.force, host(o["force"]).T)
    print(f"FP32 water48k: force Erms = {erms:.3e}")
    assert erms <= 1e-5
    assert orc.erms(r.force_cube, host(o["force_cube"]).T) <= 1e-5
    E = float(host(o["energy"])[0])
    assert abs(E - r.energy) <= 1e-5 * abs(r.energy)
    # the output sets without the virial take the two-charges-per-wave walk
    # on this grid (8^3 bins, n = 6): sphere force, +potential, cube+potential
    x32 = to_dev(xyzq, F32)
    o1 = gs.step(x32, mode=0, force=True, potential=False)
    assert orc.erms(r.force, host(o1["force"]).T) <= 1e-5
    o2 = gs.step(x32, mode=0, force=True, potential=True)
    assert orc.erms(r.force, host(o2["force"]).T) <= 1e-5
    assert rel(host(o2["pot"]), r.pot) <= 1e-5
    o3 = gs.step(x32, mode=0, force=False, potential=True, cube=True)
    assert orc.erms(r.force_cube, host(o3["force_cube"]).T) <= 1e-5
    assert rel(host(o3["pot"]), r.pot) <= 1e-5
    assert torch.equal(o2["pot"], o3["pot"])
    # (AT,CT) = (double, float): the reference's mixed instantiation
    gm = api.CpuGSEk(F64, F32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    om = gm.step(to_dev(xyzq, F32), mode=0, force=True, potential=False)
    assert orc.erms(r.force, host(om["force"]).T) <= 1e-5


def test_energy_parseval_and_real_space_agree_fp64():
    xyzq = systems.water_box(800, 30.0, seed=15)
    gs = api.CpuGSEk(F64, F64, 30, 30, 30, 2.12, 0.4, 3.0, 3.0, 30.0, 30.0, 30.0)
    o = gs.step(to_dev(xyzq), force=False, potential=False)
    Ek = float(host(o["energy"])[0])
    assert abs(gs.calculateEnergy() - Ek) <= 1e-12 * abs(Ek)


def test_corrections_closed_form():
    xyzq = np.loadtxt if False else systems.random_alternating(11, 12.0, seed=3)
    xyzq[:, 3] = np.random.default_rng(2).normal(size=11)
    gs = api.CpuGSEk(F64, F64, 12, 12, 12, 2.0, 0.4, 3.0, 3.0, 12.0, 12.0, 12.0)
    c = host(gs.corrections(to_dev(xyzq)))
    es, en = orc.corrections(2.0, (12.0,) * 3, xyzq)
    assert abs(c[0] - es) <= 1e-12 * abs(es) and abs(c[1] - en) <= 1e-12 * abs(en)


# ------------------------------------------------------ edge cases
def test_empty_system():
    gs = api.CpuGSEk(F64, F64, 16, 16, 16, 2.0, 0.4, 3.0, 3.0, 16.0, 16.0, 16.0)
    x = torch.zeros((0, 4), dtype=F64, device=DEV)
    o = gs.step(x, force=True, potential=True, virial=True)
    assert float(host(o["energy"])[0]) == 0.0
    assert np.all(host(o["virial"]) == 0.0)


def test_invalid_particles_are_rejected_and_counted():
    xyzq = systems.random_alternating(20, 16.0, seed=8)
    bad = xyzq.copy()
    bad[3, 0] = np.nan
    bad[7, 2] = np.inf
    p = orc.Params(16, 16.0, 2.0)
    g = api.CpuGaussCharge(F64, F64, 16, 16, 16)
    rho = grid(p.dims)
    g.spreadChargeToGrid(p.sigma1, p.rcut1, 20, to_dev(bad), *p.box, rho)
    assert g.invalidCount() == 2
    good = np.delete(xyzq, [3, 7], axis=0)
    assert rel(host(rho), orc.spread(p.dims, p.sigma1, p.rcut1, good, p.box)) <= TOL


def test_periodic_images_give_same_energy():
    xyzq = systems.water_box(200, 20.0, seed=9)
    shifted = xyzq.copy()
    shifted[:, :3] += np.array([20.0, -40.0, 60.0])
    gs = api.CpuGSEk(F64, F64, 20, 20, 20, 2.12, 0.4, 3.0, 3.0, 20.0, 20.0, 20.0)
    e1 = float(host(gs.step(to_dev(xyzq), force=False, potential=False)["energy"])[0])
    e2 = float(host(gs.step(to_dev(shifted), force=False, potential=False)["energy"])[0])
    assert abs(e1 - e2) <= 1e-10 * abs(e1)


def test_non_neutral_system():
    xyzq = systems.random_alternating(31, 18.0, seed=10)  # net charge -1
    p = orc.Params(18, 18.0, 2.0)
    r = orc.gse_step(p, xyzq)
    gs = api.CpuGSEk(F64, F64, 18, 18, 18, 2.0, 0.4, 3.0, 3.0, 18.0, 18.0, 18.0)
    o = gs.step(to_dev(xyzq), force=True, potential=True)
    assert abs(float(host(o["energy"])[0]) - r.energy) <= TOL * abs(r.energy)
    assert rel(host(o["force"]).T, r.force) <= TOL


def test_bad_arguments_raise():
    with pytest.raises(api._lib.GSEError):
        api.CpuGSEk(F64, F64, 16, 16, 16, 2.0, 0.7, 3.0, 3.0, 16.0, 16.0, 16.0)  # kappa > 0.5
    with pytest.raises(api._lib.GSEError):
        api.CpuGaussCharge(F32, F64, 8, 8, 8)  # (AT,CT) = (float,double) is not an instance
    g = api.CpuGaussCharge(F64, F64, 8, 8, 8)
    with pytest.raises(api._lib.GSEError):
        g.spreadChargeToGrid(-1.0, 3.0, 0, torch.zeros((0, 4), dtype=F64, device=DEV), 8.0, 8.0, 8.0,
                             grid((8, 8, 8)))


def test_overflow_cluster_parity_and_determinism():
    """A dense cluster overfills its bin (capacity 2*mean + 32): the excess
    goes through the overflow path.  FP64 stays within 1e-10 of the oracle;
    FP32 grids are bit-identical across reruns and particle permutations,
    whichever particles land on the overflow list."""
    rng = np.random.default_rng(17)
    n = 2000
    xyzq = np.empty((n, 4))
    xyzq[:, :3] = 10.0 + rng.random((n, 3)) * 0.9  # one cell of a 32^3 grid, h = 0.97
    xyzq[:, 3] = 0.25 + 0.5 * rng.random(n)  # one sign: no cancellation in the grid
    p = orc.Params(32, 31.0, 2.12)
    ref = orc.gse_step(p, xyzq)
    g64 = api.CpuGSEk(F64, F64, 32, 32, 32, 2.12, 0.4, 3.0, 3.0, 31.0, 31.0, 31.0)
    o = g64.step(to_dev(xyzq), force=True, potential=True)
    assert rel(host(o["force"]).T, ref.force) <= TOL
    assert abs(float(host(o["energy"])[0]) - ref.energy) / abs(ref.energy) <= TOL
    g32 = api.CpuGSEk(F32, F32, 32, 32, 32, 2.12, 0.4, 3.0, 3.0, 31.0, 31.0, 31.0)
    x32 = to_dev(xyzq, F32)
    g32.spreadCharge1(n, x32)
    r1 = g32.getRhoS().clone()
    for seed in range(3):
        perm = torch.randperm(n, generator=torch.Generator().manual_seed(seed)).to(DEV)
        g32.spreadCharge1(n, x32[perm].contiguous())
        assert torch.equal(g32.getRhoS(), r1)
    assert rel(host(r1), orc.spread(p.dims, p.sigma1, p.rcut1, host(x32), p.box)) <= 1e-6


@pytest.mark.parametrize("R", [5, 3, 7])
def test_convolution_fp32_asymmetric_taps(R):
    """Taps that are not symmetric (no pairing) and radii other than 5 (no
    register window) still match the oracle's three-pass convolution."""
    dims = (24, 20, 28)
    X, Y, Z = dims
    rng = np.random.default_rng(20 + R)
    rho = rng.normal(size=(Z, Y, X))
    ks = [rng.random(2 * R + 1) for _ in range(3)]
    c = api.CpuConvolution(F32, *dims)
    c.setupKernel(R, R, R, *ks)
    o = grid(dims, F32)
    c.runXYZ(to_dev(rho, F32), o)
    k32 = [k.astype(np.float32).astype(np.float64) for k in ks]
    ref = orc.conv_run_xyz(dims, *k32, rho.astype(np.float32).astype(np.float64))
    assert rel(host(o), ref) < 1e-6
    c.run(to_dev(rho, F32), o)  # X pass only (CpuConvolution::run)
    ref_x = orc.conv_run_x(dims, k32[0], rho.astype(np.float32).astype(np.float64))
    assert rel(host(o), ref_x) < 1e-6


@pytest.mark.parametrize("dims,dt,sym", [
    ((9, 4, 7), F64, False),      # axes shorter than the radius: repeated wraps
    ((9, 4, 7), F32, True),
    ((40, 33, 50), F32, False),   # ragged lengths, last tile partly outside
    ((256, 40, 256), F32, True),  # y pass in 32-row tiles, z pass in 16-row tiles
    ((256, 40, 256), F64, False),
    ((8, 5, 6), F32, True),       # shortest row of the 4-wide x pass (halo from 8 values)
    ((12, 9, 10), F64, False),
    ((256, 130, 256), F32, True),  # y pass in 32-row tiles, last tile 2 rows
    ((256, 130, 256), F64, False),
])
def test_convolution_pencil_shapes(dims, dt, sym):
    """Radius-5 passes (4-wide x pass, register-pencil y and z passes):
    periodic wraps, ragged tiles and both tile heights match the oracle's
    three-pass convolution."""
    X, Y, Z = dims
    R = 5
    rng = np.random.default_rng(X * 7 + Y + Z)
    rho = rng.normal(size=(Z, Y, X))
    if sym:
        k = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * 2.0 ** 2))
        ks = [k, k * 0.5, k * 1.5]
    else:
        ks = [rng.random(2 * R + 1) for _ in range(3)]
    c = api.CpuConvolution(dt, *dims)
    c.setupKernel(R, R, R, *ks)
    o = grid(dims, dt)
    c.runXYZ(to_dev(rho, dt), o)
    if dt == F32:
        k32 = [k.astype(np.float32).astype(np.float64) for k in ks]
        ref = orc.conv_run_xyz(dims, *k32, rho.astype(np.float32).astype(np.float64))
        assert rel(host(o), ref) < 1e-6
    else:
        assert rel(host(o), orc.conv_run_xyz(dims, *ks, rho)) <= TOL


@pytest.mark.parametrize("dt", [F32, F64])
def test_binning_variants_give_identical_grids(dt):
    """Every binning variant (one slot claim per charge, or claims aggregated
    per workgroup of 256 to 2048 charges) gives bit-identical grids and
    energies: a water box with a dense cluster that overfills its bin and two
    rejected charges, n not a multiple of any workgroup size.  Which cluster
    charges take the overflow path differs between variants, so per-charge
    outputs are checked against the oracle (FP64 1e-10, FP32 Erms 1e-5)."""
    rng = np.random.default_rng(23)
    xyzq = systems.water_box(3001, 45.0, seed=23)
    cl = np.empty((700, 4))
    cl[:, :3] = 20.0 + rng.random((700, 3)) * 0.9
    cl[:, 3] = 0.1 * rng.standard_normal(700)
    xyzq = np.concatenate([xyzq, cl])
    xyzq[11, 1] = np.nan
    xyzq[5000, 3] = np.inf
    M = 48
    p = orc.Params(M, 45.0, 2.12)
    good = np.delete(xyzq, [11, 5000], axis=0)
    if dt == F32:
        good = good.astype(np.float32).astype(np.float64)
    ref = orc.gse_step(p, good)
    keep = np.ones(len(xyzq), bool)
    keep[[11, 5000]] = False
    keys = api.get_tuning()
    try:
        outs = []
        for agg in (0, 1, 2, 4, 8):
            api.set_tuning(bin_agg=agg)
            g = api.CpuGaussCharge(dt, dt, M, M, M)
            rho = grid(p.dims, dt)
            g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq, dt), *p.box, rho)
            assert g.invalidCount() == 2
            gs = api.CpuGSEk(dt, dt, M, M, M, 2.12, 0.4, 3.0, 3.0, 45.0, 45.0, 45.0)
            o = gs.step(to_dev(xyzq, dt), force=True, potential=True)
            f = host(o["force"]).T[keep]
            if dt == F64:
                assert rel(f, ref.force) <= TOL
                assert rel(host(rho), orc.spread(p.dims, p.sigma1, p.rcut1, good, p.box)) <= TOL
            else:
                assert orc.erms(ref.force, f) <= 1e-5
            outs.append((rho, o["energy"].clone()))
        for r in outs[1:]:
            for a, b in zip(outs[0], r):
                assert torch.equal(a, b)
    finally:
        api.set_tuning(bin_agg=keys["bin_agg"])


@pytest.mark.parametrize("dt", [F32, F64])
@pytest.mark.parametrize("ncluster", [600, 3000], ids=["bin_overflow", "superbin_overflow"])
def test_superbin_binning_above_16384_bins(dt, ncluster):
    """Grids of more than 16384 bins are binned in two passes (super-bins of
    4^3 bins, then the bins of each super-bin).  It must give the same
    results as one slot claim per charge, bit for bit: grids, energy and the
    per-charge outputs, with a cluster that overfills its bin (600 charges)
    or its whole super-bin (3000, past the super-bin's 2x mean + 256 slots)
    and two rejected charges (M = 216: 19683 bins of 8 cells).  Which charges
    of the cluster's super-bin take the overflow path (gathered from global
    memory, summed in another order) differs between the two binnings, so
    their outputs agree to rounding only."""
    rng = np.random.default_rng(31)
    L, M, n = 215.3, 216, 120000
    xyzq = np.empty((n, 4))
    xyzq[:, :3] = rng.random((n, 3)) * L
    xyzq[:, 3] = np.where(np.arange(n) % 2 == 0, 0.5, -0.5)
    xyzq[:ncluster, :3] = 50.0 + rng.random((ncluster, 3)) * 0.9   # overfills one bin
    xyzq[7, 0] = np.nan
    xyzq[9, 3] = np.inf
    x = to_dev(xyzq, dt)
    keys = api.get_tuning()
    outs = []
    try:
        for agg in (4, 0):
            api.set_tuning(bin_agg=agg)
            gs = api.CpuGSEk(dt, dt, M, M, M, 2.12, 0.4, 3.0, 3.0, L, L, L)
            o = gs.step(x, force=True, potential=True)
            lay = gs.layout()
            assert lay["nbins"] == 19683
            outs.append((gs.getRhoS().clone(), o["energy"].clone(), o["force"].clone(),
                         o["pot"].clone()))
    finally:
        api.set_tuning(bin_agg=keys["bin_agg"])
    # every charge of the cluster's super-bin may take either path; the rest
    # must agree bit for bit
    good = np.isfinite(xyzq).all(axis=1)
    o = np.mod(np.floor(np.nan_to_num(xyzq[:, :3]) / (L / M) + 0.5).astype(np.int64) - 6, M) // 32
    sup = o[:, 0] + 7 * (o[:, 1] + 7 * o[:, 2])
    hot = np.isin(sup, sup[10:ncluster])
    keep_np = good & ~hot
    keep = torch.from_numpy(keep_np).to(DEV)
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert torch.equal(outs[0][2][:, keep], outs[1][2][:, keep])
    assert torch.equal(outs[0][3][keep], outs[1][3][keep])
    cl = torch.from_numpy(good & hot).to(DEV)
    tol = 1e-5 if dt == F32 else 1e-12
    assert rel(host(outs[0][2][:, cl]), host(outs[1][2][:, cl])) <= tol
    assert rel(host(outs[0][3][cl]), host(outs[1][3][cl])) <= tol


@pytest.mark.parametrize("dt", [F32, F64])
@pytest.mark.parametrize("wl", ["water_48k_M80", "superbin"])
def test_bin_order_gives_identical_results(dt, wl):
    """The spread's and the gather's workgroups take each XCD's bins by
    descending charge count of the previous binning (bin_order=1) or in bin
    order (0).  Only the schedule changes: grids, energy, forces and
    potentials agree bit for
    bit, on the 48k box (1000 bins, 17 to 92 charges each) and on a 216^3
    grid (19683 bins, super-bin binning) with a cluster that overfills one
    bin.  Which charges of that cluster's super-bin take the overflow path
    (gathered from global memory, summed in another order) depends on the
    order of the binning's slot claims, so their outputs agree to rounding
    only."""
    if wl == "superbin":
        rng = np.random.default_rng(5)
        L, M, sigma, n = 215.3, 216, 2.12, 60000
        xyzq = np.empty((n, 4))
        xyzq[:, :3] = rng.random((n, 3)) * L
        xyzq[:, 3] = np.where(np.arange(n) % 2 == 0, 0.5, -0.5)
        xyzq[:400, :3] = 80.0 + rng.random((400, 3)) * 0.9
    else:
        xyzq, L, M, sigma = systems.make(wl)
    x = to_dev(xyzq, dt)
    old = api.get_tuning()["bin_order"]
    outs = []
    try:
        for bo in (1, 0):
            api.set_tuning(bin_order=bo)
            gs = api.CpuGSEk(dt, dt, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
            # the first step runs in bin order and sorts its counts; the
            # second takes that schedule
            for _ in range(2):
                o = gs.step(x, force=True, potential=True, virial=True)
            outs.append([gs.getRhoS().clone()] + [o[k].clone() for k in
                                                 ("energy", "force", "pot", "virial")])
    finally:
        api.set_tuning(bin_order=old)
    keep_np = np.ones(len(xyzq), dtype=bool)
    if wl == "superbin":
        o = np.mod(np.floor(xyzq[:, :3] / (L / M) + 0.5).astype(np.int64) - 6, M) // 32
        sup = o[:, 0] + 7 * (o[:, 1] + 7 * o[:, 2])
        keep_np = ~np.isin(sup, sup[:400])
    keep = torch.from_numpy(keep_np).to(DEV)
    for a, b in zip(outs[0][:2], outs[1][:2]):  # grid and energy
        assert torch.equal(a, b)
    assert torch.equal(outs[0][2][:, keep], outs[1][2][:, keep])
    assert torch.equal(outs[0][3][keep], outs[1][3][keep])
    if wl == "superbin":
        tol = 1e-5 if dt == F32 else 1e-12
        hot = torch.from_numpy(~keep_np).to(DEV)
        assert rel(host(outs[0][2][:, hot]), host(outs[1][2][:, hot])) <= tol
        assert rel(host(outs[0][3][hot]), host(outs[1][3][hot])) <= tol
        assert rel(host(outs[0][4]), host(outs[1][4])) <= tol
    else:
        assert torch.equal(outs[0][4], outs[1][4])


def _tiling_system(seed):
    """A 48^3 grid in a 46 A box: h = 0.958, n = 6, so 8^3 bins with 20^3
    bricks (64-bit bricks fit the LDS budget) and 6 bins per axis: every
    tiled reduce applies.  A dense cluster overfills one bin, so the
    reduce also adds the overflow particles' int64 grid."""
    rng = np.random.default_rng(seed)
    xyzq = systems.water_box(2000, 46.0, seed=seed)
    cl = np.empty((500, 4))
    cl[:, :3] = 21.0 + rng.random((500, 3)) * 0.9
    cl[:, 3] = 0.1 * rng.standard_normal(500)
    return np.concatenate([xyzq, cl]), 48, 46.0


def _overflowed(layout, xyzq, p):
    """True when some bin received more particles than its slots (checked on
    the host from the cube origins, as the binning kernel assigns them)."""
    B = layout["bin"][0]
    o = np.floor(xyzq[:, :3] / (p.box[0] / p.X) + 0.5).astype(np.int64) - 6
    o = np.mod(o, p.X) // B
    nb = p.X // B
    _, cnt = np.unique(o[:, 0] + nb * (o[:, 1] + nb * o[:, 2]), return_counts=True)
    return cnt.max() > layout["cap_bin"]


def test_reduce_load_modes_give_identical_grids():
    """The FP32 tiled brick reduce with ordinary or non-temporal brick loads
    sums the same integers, so the grids and energies are bit-identical;
    forces match the FP64 oracle to Erms 1e-5.  The layout getter confirms
    that each mode ran its own kernel, with the overflow grid in use."""
    xyzq, M, L = _tiling_system(29)
    p = orc.Params(M, L, 2.12)
    ref = orc.gse_step(p, xyzq.astype(np.float32).astype(np.float64))
    keys = api.get_tuning()
    try:
        outs = []
        for mode, path in ((0, "tile4"), (1, "tile4_nt")):
            api.set_tuning(reduce_nt=mode)
            g = api.CpuGaussCharge(F32, F32, M, M, M)
            rho = grid(p.dims, F32)
            g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq, F32), *p.box, rho)
            lay = g.layout()
            assert lay["reduce"] == path, lay
            assert lay["bin"] == (8, 8, 8) and lay["brick"] == (20, 20, 20), lay
            assert _overflowed(lay, xyzq, p)
            gs = api.CpuGSEk(F32, F32, M, M, M, 2.12, 0.4, 3.0, 3.0, L, L, L)
            o = gs.step(to_dev(xyzq, F32), force=True, potential=True)
            assert gs.layout()["reduce"] == path
            assert orc.erms(ref.force, host(o["force"]).T) <= 1e-5
            outs.append((rho, o["energy"].clone()))
        for r in outs[1:]:
            for a, b in zip(outs[0], r):
                assert torch.equal(a, b)
    finally:
        api.set_tuning(reduce_nt=keys["reduce_nt"])


@pytest.mark.parametrize("dt,lds32", [(F64, 1), (F32, 0)], ids=["F64", "F32-int64"])
def test_reduce_tile2_gives_identical_grids(dt, lds32):
    """The tiled reduce of 64-bit bricks with two x points per thread and the
    one-point form sum the same integers: grids and energies bit-identical,
    FP64 grid and forces within 1e-10 of the oracle (FP32: Erms 1e-5).  The
    layout getter confirms each kernel ran, with the overflow grid in use."""
    xyzq, M, L = _tiling_system(31)
    p = orc.Params(M, L, 2.12)
    src = xyzq if dt == F64 else xyzq.astype(np.float32).astype(np.float64)
    ref = orc.gse_step(p, src)
    rho_ref = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq, p.box) if dt == F64 else None
    keys = api.get_tuning()
    try:
        api.set_tuning(spread_lds32=lds32)
        outs = []
        for mode, path in ((1, "tile2"), (0, "tile")):
            api.set_tuning(reduce_tile2=mode)
            g = api.CpuGaussCharge(dt, dt, M, M, M)
            rho = grid(p.dims, dt)
            g.spreadChargeToGrid(p.sigma1, p.rcut1, len(xyzq), to_dev(xyzq, dt), *p.box, rho)
            lay = g.layout()
            assert lay["reduce"] == path, lay
            assert _overflowed(lay, xyzq, p)
            gs = api.CpuGSEk(dt, dt, M, M, M, 2.12, 0.4, 3.0, 3.0, L, L, L)
            o = gs.step(to_dev(xyzq, dt), force=True, potential=True)
            assert gs.layout()["reduce"] == path
            f = host(o["force"]).T
            if dt == F64:
                assert rel(host(rho), rho_ref) <= TOL
                assert rel(f, ref.force) <= TOL
            else:
                assert orc.erms(ref.force, f) <= 1e-5
            outs.append((rho, o["energy"].clone()))
        for a, b in zip(outs[0], outs[1]):
            assert torch.equal(a, b)
    finally:
        api.set_tuning(reduce_tile2=keys["reduce_tile2"], spread_lds32=keys["spread_lds32"])


def test_numcoord_beyond_xyzq_is_rejected():
    """numCoord larger than the particle tensor is refused on the host, before
    any kernel could read past the allocation."""
    x = to_dev(systems.water_box(10, 16.0, seed=2))
    g = api.CpuGaussCharge(F64, F64, 16, 16, 16)
    rho = grid((16, 16, 16))
    with pytest.raises(ValueError):
        g.spreadChargeToGrid(1.34, 5.69, 31, x, 16.0, 16.0, 16.0, rho)
    f = torch.zeros(40, dtype=F64, device=DEV)
    with pytest.raises(ValueError):
        g.interpolateForce(1.34, 5.69, 31, x, 16.0, 16.0, 16.0, rho, f, f, f)
    gs = api.CpuGSEk(F64, F64, 16, 16, 16, 2.12, 0.4, 3.0, 3.0, 16.0, 16.0, 16.0)
    with pytest.raises(ValueError):
        gs.spreadCharge1(-1, x)
    with pytest.raises(ValueError):
        gs.interpolateForce(31, x, f, f, f)
