"""Pins the FP64 oracle (oracle/gse_oracle.c) before anything is checked
against it.

Each oracle routine is checked by something that would fail on a plausible
mistake in it (dropped term, wrong sign, transposed index, off-by-one cube):
  * the reference's own compiled CPU classes (oracle/_ref, bit-for-bit),
  * the committed golden fixtures written from oracle/_ref,
  * independent numpy brute force (explicit image sums, direct DFT),
  * numpy.fft (pocketfft) for the 3D transforms that stand in for FFTW,
  * closed forms and invariants (Parseval, single Fourier modes, Gaussian
    sum rules, linearity, finite differences of the oracle's own energy).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from gse_amd import systems

dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(dp)


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b)
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


# ---------------------------------------------------------------- brute force
def brute_spread(dims, sigma, rcut, xyzq, box):
    """rho_p = sum_i q_i (2 pi s^2)^-3/2 sum over the (2n+1)^3 cube centred
    on round(x/h) of exp(-|r_i - t h|^2 / 2 s^2), cube images wrapped."""
    X, Y, Z = dims
    h = np.array(box, dtype=float) / np.array(dims)
    n = [int(math.ceil(rcut / hh)) for hh in h]
    pref = (2 * np.pi * sigma * sigma) ** -1.5
    rho = np.zeros((Z, Y, X))
    for x, y, z, q in xyzq:
        c = [int(round(v / hh)) for v, hh in zip((x, y, z), h)]
        tx = np.arange(c[0] - n[0], c[0] + n[0] + 1)
        ty = np.arange(c[1] - n[1], c[1] + n[1] + 1)
        tz = np.arange(c[2] - n[2], c[2] + n[2] + 1)
        wx = np.exp(-(x - tx * h[0]) ** 2 / (2 * sigma ** 2))
        wy = np.exp(-(y - ty * h[1]) ** 2 / (2 * sigma ** 2))
        wz = np.exp(-(z - tz * h[2]) ** 2 / (2 * sigma ** 2))
        w = q * pref * wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        np.add.at(rho, (np.mod(tz, Z)[:, None, None], np.mod(ty, Y)[None, :, None],
                        np.mod(tx, X)[None, None, :]), w)
    return rho


def brute_gather(dims, sigma, rcut, xyzq, box, phi, sphere=True):
    X, Y, Z = dims
    h = np.array(box, dtype=float) / np.array(dims)
    n = [int(math.ceil(rcut / hh)) for hh in h]
    prefF = (2 * np.pi) ** -1.5 * h.prod() * sigma ** -5
    out = np.zeros((len(xyzq), 3))
    for i, (x, y, z, q) in enumerate(xyzq):
        c = [int(round(v / hh)) for v, hh in zip((x, y, z), h)]
        tx = np.arange(c[0] - n[0], c[0] + n[0] + 1)
        ty = np.arange(c[1] - n[1], c[1] + n[1] + 1)
        tz = np.arange(c[2] - n[2], c[2] + n[2] + 1)
        dz, dy, dx = np.meshgrid(z - tz * h[2], y - ty * h[1], x - tx * h[0], indexing="ij")
        r2 = dx * dx + dy * dy + dz * dz
        ph = phi[np.ix_(np.mod(tz, Z), np.mod(ty, Y), np.mod(tx, X))]
        s = q * prefF * np.exp(-r2 / (2 * sigma ** 2)) * ph
        if sphere:
            s = np.where(r2 < rcut * rcut, s, 0.0)
        out[i] = [np.sum(s * dx), np.sum(s * dy), np.sum(s * dz)]
    return out


def dft3(a, sign):
    """O(N^2) direct 3D DFT of a[z][y][x]."""
    Z, Y, X = a.shape
    out = a.astype(np.complex128)
    for ax, m in ((2, X), (1, Y), (0, Z)):
        k = np.arange(m)
        W = np.exp(sign * 2j * np.pi * np.outer(k, k) / m)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 20, 24, 25, 27, 30, 32, 49, 80])
def test_fft1d_vs_direct_dft(orc, n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    k = np.arange(n)
    for sign in (-1, 1):
        ref = np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ a
        assert rel(orc.fft_1d(a, sign), ref) < 1e-13


@pytest.mark.parametrize("dims", [(4, 6, 5), (8, 8, 8), (9, 7, 6), (20, 20, 20), (32, 24, 40), (5, 3, 4)])
def test_fft3d_vs_numpy(orc, dims):
    X, Y, Z = dims
    rng = np.random.default_rng(sum(dims))
    g = rng.normal(size=(Z, Y, X))
    spec = orc.fft_r2c(g)
    assert rel(spec, np.fft.rfftn(g)) < 1e-13
    # unnormalised inverse: numpy's irfftn divides by the size
    back = orc.fft_c2r(spec, X)
    assert rel(back, np.fft.irfftn(spec, s=(Z, Y, X), axes=(0, 1, 2)) * X * Y * Z) < 1e-13
    assert rel(back, g * X * Y * Z) < 1e-13


def test_fft3d_vs_direct_dft(orc):
    rng = np.random.default_rng(7)
    g = rng.normal(size=(5, 3, 4))
    assert rel(orc.fft_r2c(g), dft3(g, -1)[:, :, :3]) < 1e-13


# ----------------------------------------------------------------- Green's
@pytest.mark.parametrize("dims,box", [((8, 8, 8), (10.0, 10.0, 10.0)), ((9, 6, 7), (11.0, 8.5, 9.5))])
@pytest.mark.parametrize("kind,sigma", [(0, 0.9486), (0, 0.0), (1, 0.9486)])
def test_greens_mul_vs_reference(orc, ref, dims, box, kind, sigma):
    X, Y, Z = dims
    rng = np.random.default_rng(3)
    spec = rng.normal(size=(Z, Y, X // 2 + 1)) + 1j * rng.normal(size=(Z, Y, X // 2 + 1))
    mine = orc.greens_mul(kind, dims, box, sigma, spec)
    theirs = np.ascontiguousarray(spec.copy())
    ref.ref_greens_mul_d(kind, X, Y, Z, *map(ctypes.c_double, box), ctypes.c_double(sigma),
                         theirs.ctypes.data_as(dp))
    assert np.array_equal(mine, theirs)


def test_greens_single_mode(orc):
    """rho = cos(2 pi m.p / M) gives phi = M^3 G(k_m) rho (unnormalised FFTs)."""
    dims = (12, 10, 8)
    box = (13.0, 11.0, 9.0)
    X, Y, Z = dims
    m = (2, -3, 1)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    rho = np.cos(2 * np.pi * (m[0] * xx / X + m[1] * yy / Y + m[2] * zz / Z))
    for sigma in (0.0, 0.95):
        phi = orc.fft_solve(0, dims, sigma, box, rho)
        k = 2 * np.pi * np.array(m) / np.array(box)
        k2 = k @ k
        G = 4 * np.pi / (X * Y * Z) * np.exp(-sigma ** 2 / 2 * k2) / k2
        assert rel(phi, X * Y * Z * G * rho) < 1e-12
    # the k = 0 mode is removed
    phi0 = orc.fft_solve(0, dims, 0.95, box, np.ones((Z, Y, X)))
    assert np.max(np.abs(phi0)) < 1e-12


def test_nyquist_and_odd_index_mapping(orc):
    """Index ix >= ceil(M/2) maps to ix - M: for even M the Nyquist plane
    is -M/2, for odd M the largest positive index is (M-1)/2."""
    for dims in ((6, 6, 6), (7, 5, 9)):
        X, Y, Z = dims
        box = (7.0, 6.0, 8.0)
        spec = np.ones((Z, Y, X // 2 + 1), dtype=np.complex128)
        out = orc.greens_mul(1, dims, box, 1.3, spec)
        for iz in range(Z):
            for iy in range(Y):
                for ix in range(X // 2 + 1):
                    kx = ix - X if ix >= (X + 1) // 2 else ix
                    ky = iy - Y if iy >= (Y + 1) // 2 else iy
                    kz = iz - Z if iz >= (Z + 1) // 2 else iz
                    k2 = sum((2 * np.pi * kk / b) ** 2 for kk, b in zip((kx, ky, kz), box))
                    assert abs(out[iz, iy, ix] - np.exp(-1.3 ** 2 / 2 * k2) / (X * Y * Z)) < 1e-16


# ------------------------------------------------------------------ spread
CASES = [
    # dims, box, sigma, n particles
    ((20, 20, 20), (20.0, 20.0, 20.0), 1.341640787, "pair"),
    ((16, 12, 14), (15.0, 11.0, 13.5), 1.2, "random"),
    ((8, 8, 8), (8.0, 8.0, 8.0), 1.34, "random"),   # 2n+1 = 13 > M: images wrap
    ((32, 32, 32), (31.0, 31.0, 31.0), 1.3416, "water"),
]


def case_particles(kind, box, seed=11):
    if kind == "pair":
        return systems.pair(1.0, box[0], orientation="diagonal")
    if kind == "water":
        return systems.water_box(40, box[0], seed=seed)
    rng = np.random.default_rng(seed)
    a = np.empty((25, 4))
    a[:, :3] = rng.random((25, 3)) * np.array(box)
    a[:, 3] = rng.normal(size=25)
    a[0, :3] = [-1.3, -0.2, -2.0]  # negative coordinates are legal
    return a


@pytest.mark.parametrize("dims,box,sigma,kind", CASES)
@pytest.mark.parametrize("normalize", [False, True])
def test_spread_vs_reference_bitexact(orc, ref, dims, box, sigma, kind, normalize):
    xyzq = case_particles(kind, box)
    rcut = 3 * math.sqrt(2) * sigma
    mine = orc.spread(dims, sigma, rcut, xyzq, box, normalize)
    X, Y, Z = dims
    theirs = np.zeros((Z, Y, X))
    a = np.ascontiguousarray(xyzq)
    ref.ref_spread_dd(X, Y, Z, ctypes.c_double(sigma), ctypes.c_double(rcut), len(a), P(a),
                      *map(ctypes.c_double, box), P(theirs), int(normalize))
    assert np.array_equal(mine, theirs)


@pytest.mark.parametrize("dims,box,sigma,kind", CASES)
def test_spread_vs_brute_force(orc, dims, box, sigma, kind):
    xyzq = case_particles(kind, box)
    rcut = 3 * math.sqrt(2) * sigma
    mine = orc.spread(dims, sigma, rcut, xyzq, box)
    assert rel(mine, brute_spread(dims, sigma, rcut, xyzq, box)) < 1e-13


def test_spread_sum_rule_and_linearity(orc):
    dims, box, sigma = (16, 16, 16), (16.5, 16.5, 16.5), 1.3
    rcut = 3 * math.sqrt(2) * sigma
    xyzq = case_particles("random", box, seed=5)
    rho = orc.spread(dims, sigma, rcut, xyzq, box)
    h = np.array(box) / np.array(dims)
    # sum_p rho_p h^3 = sum_i q_i prod_d (h_d sum_j w_d[j]) (2 pi s^2)^-1/2
    tot = 0.0
    for x, y, z, q in xyzq:
        f = q
        for v, hh in zip((x, y, z), h):
            nn = int(math.ceil(rcut / hh))
            c = int(round(v / hh))
            t = np.arange(c - nn, c + nn + 1)
            f *= hh * np.sum(np.exp(-(v - t * hh) ** 2 / (2 * sigma ** 2))) / math.sqrt(2 * np.pi * sigma ** 2)
        tot += f
    assert abs(rho.sum() * h.prod() - tot) < 1e-12 * np.abs(xyzq[:, 3]).sum()
    # linearity in q
    x2 = xyzq.copy()
    x2[:, 3] *= -2.5
    assert rel(orc.spread(dims, sigma, rcut, x2, box), -2.5 * rho) < 1e-14
    # moving one particle by exactly one grid spacing shifts its image by one index
    one = xyzq[:1].copy()
    r1 = orc.spread(dims, sigma, rcut, one, box)
    one[0, 1] += h[1]
    r2 = orc.spread(dims, sigma, rcut, one, box)
    assert rel(r2, np.roll(r1, 1, axis=1)) < 1e-12


def test_support_size_vs_reference(orc, ref):
    for dims, box, rcut in (((20, 20, 20), (20.0,) * 3, 5.6921), ((32, 24, 16), (31.0, 25.0, 17.0), 4.02)):
        assert orc.support_size(dims, rcut, box) == ref.ref_support_size(*dims, ctypes.c_double(rcut),
                                                                         *map(ctypes.c_double, box))


# ----------------------------------------------------------------- gather
@pytest.mark.parametrize("dims,box,sigma,kind", CASES)
def test_interpolate_vs_reference_bitexact(orc, ref, dims, box, sigma, kind):
    xyzq = case_particles(kind, box)
    rcut = 3 * math.sqrt(2) * sigma
    X, Y, Z = dims
    rng = np.random.default_rng(9)
    phi = rng.normal(size=(Z, Y, X))
    mine = orc.interpolate(dims, sigma, rcut, xyzq, box, phi)
    a = np.ascontiguousarray(xyzq)
    f = np.zeros((3, len(a)))
    ref.ref_interpolate_dd(X, Y, Z, ctypes.c_double(sigma), ctypes.c_double(rcut), len(a), P(a),
                           *map(ctypes.c_double, box), P(phi), P(f[0]), P(f[1]), P(f[2]))
    assert np.array_equal(mine, f.T)


@pytest.mark.parametrize("dims,box,sigma,kind", CASES[:3])
def test_interpolate_vs_brute_force(orc, dims, box, sigma, kind):
    xyzq = case_particles(kind, box)
    rcut = 3 * math.sqrt(2) * sigma
    X, Y, Z = dims
    phi = np.random.default_rng(9).normal(size=(Z, Y, X))
    mine = orc.interpolate(dims, sigma, rcut, xyzq, box, phi)
    assert rel(mine, brute_gather(dims, sigma, rcut, xyzq, box, phi, sphere=True)) < 1e-12
    _, fc, _ = orc.gather_cube(dims, sigma, rcut, xyzq, box, phi)
    assert rel(fc, brute_gather(dims, sigma, rcut, xyzq, box, phi, sphere=False)) < 1e-12
    # the sphere and cube forces differ only by the corner terms, which are small
    assert 0 < rel(mine, fc) < 1e-2


# ----------------------------------------------------------- convolution
def test_spread_on_grid_vs_reference_and_brute(orc, ref):
    dims, box, sigma = (12, 10, 14), (12.5, 10.0, 13.0), 0.9487
    rcut = 3 * math.sqrt(2) * sigma
    X, Y, Z = dims
    rho = np.random.default_rng(4).normal(size=(Z, Y, X))
    mine = orc.spread_on_grid(dims, sigma, rcut, box, rho)
    theirs = np.zeros_like(rho)
    ref.ref_spread_on_grid_dd(X, Y, Z, ctypes.c_double(sigma), ctypes.c_double(rcut),
                              *map(ctypes.c_double, box), P(rho), P(theirs))
    assert np.array_equal(mine, theirs)
    # brute force 3D periodic convolution with the separable kernel
    h = np.array(box) / np.array(dims)
    ns = [int(math.ceil(rcut / hh)) for hh in h]
    pref = (2 * np.pi * sigma ** 2) ** -1.5 * h.prod()
    out = np.zeros_like(rho)
    for dz in range(-ns[2], ns[2] + 1):
        for dy in range(-ns[1], ns[1] + 1):
            for dx in range(-ns[0], ns[0] + 1):
                w = pref * np.exp(-((dx * h[0]) ** 2 + (dy * h[1]) ** 2 + (dz * h[2]) ** 2) / (2 * sigma ** 2))
                out += w * np.roll(rho, (-dz, -dy, -dx), axis=(0, 1, 2))
    assert rel(mine, out) < 1e-13


def test_conv_run_x_vs_reference(orc, ref):
    dims = (16, 8, 6)
    X, Y, Z = dims
    rng = np.random.default_rng(12301)
    g = rng.normal(size=(Z, Y, X))
    R = 5
    kx = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * 2.12 ** 2))
    mine = orc.conv_run_x(dims, kx, g)
    theirs = np.zeros_like(g)
    ref.ref_conv_run_d(X, Y, Z, R, R, R, P(kx), P(kx), P(kx), P(g), P(theirs))
    assert np.array_equal(mine, theirs)
    # X only: each row is a circular convolution
    row = np.array([sum(kx[j + R] * g[1, 2, (i + j) % X] for j in range(-R, R + 1)) for i in range(X)])
    assert rel(mine[1, 2], row) < 1e-14


def test_charge_spread_identity(orc):
    """sigma1 spread then sigma2 convolution ~ one spread with
    sigma_t = sqrt(sigma1^2 + sigma2^2): the reference's TEST_CHARGE_SPREAD
    check (cpu_gse.cpp:752-794), real-space and k-space (G2) routes."""
    p = orc.Params(20, 20.0, 2.121320344)
    xyzq = systems.pair(1.0, 20.0, orientation="diagonal")
    rho1 = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq, p.box)
    st = math.sqrt(p.sigma1 ** 2 + p.sigma2 ** 2)
    ref_t = orc.spread(p.dims, st, 3 * math.sqrt(2) * st, xyzq, p.box)
    real = orc.spread_on_grid(p.dims, p.sigma2, p.rcut2, p.box, rho1)
    kspace = orc.fft_solve(1, p.dims, p.sigma2, p.box, rho1)
    assert rel(real, ref_t) < 2e-3
    assert rel(kspace, ref_t) < 2e-3


# ------------------------------------------------------------- energies
def test_parseval_energy_and_gse_step(orc):
    p = orc.Params(32, 31.0, 2.12)
    xyzq = systems.water_box(60, 31.0, seed=8)
    r = orc.gse_step(p, xyzq, mode=0, want_grids=True)
    assert abs(r.energy - r.energy_k) < 1e-12 * abs(r.energy)
    # E = 1/2 sum_i q_i phi_i exactly (cube support on both sides)
    assert abs(0.5 * np.dot(xyzq[:, 3], r.pot) - r.energy) < 1e-11 * abs(r.energy)
    # composition of the individual routines
    rho = orc.spread(p.dims, p.sigma1, p.rcut1, xyzq, p.box)
    phi = orc.fft_solve(0, p.dims, p.sigma2, p.box, rho)
    assert np.array_equal(rho, r.rho)
    assert np.array_equal(phi, r.phi)
    assert orc.energy(p.dims, p.ccelec, p.box, rho, phi) == r.energy
    f = orc.interpolate(p.dims, p.sigma1, p.rcut1, xyzq, p.box, phi) * p.ccelec
    assert np.array_equal(f, r.force)


def gse_energy(orc, p, xyzq, mode=0):
    return orc.gse_step(p, xyzq, mode=mode).energy


def safe_particles(n, L, h, seed):
    """Random particles kept away from cube-selection boundaries (x/h within
    0.05 of a half-integer), where the energy jumps."""
    rng = np.random.default_rng(seed)
    a = np.empty((n, 4))
    for i in range(n):
        while True:
            v = rng.random(3) * L
            f = np.abs(np.mod(v / h, 1.0) - 0.5)
            if np.all(f > 0.05):
                break
        a[i, :3] = v
    a[:, 3] = rng.normal(size=n)
    a[:, 3] -= a[:, 3].mean()
    return a


def test_cube_force_is_minus_energy_gradient(orc):
    p = orc.Params(16, 16.0, 2.0)
    xyzq = safe_particles(6, 16.0, 1.0, 21)
    r = orc.gse_step(p, xyzq)
    eps = 1e-5
    for i in (0, 3):
        for d in range(3):
            a = xyzq.copy()
            a[i, d] += eps
            ep = gse_energy(orc, p, a)
            a[i, d] -= 2 * eps
            em = gse_energy(orc, p, a)
            fd = -(ep - em) / (2 * eps)
            assert abs(fd - r.force_cube[i, d]) < 1e-6 * np.max(np.abs(r.force_cube))
    # the reference (sphere) force is close to, but not exactly, the gradient
    assert 1e-9 < rel(r.force, r.force_cube) < 1e-2


def test_potential_is_energy_charge_derivative(orc):
    p = orc.Params(16, 16.0, 2.0)
    xyzq = safe_particles(5, 16.0, 1.0, 22)
    r = orc.gse_step(p, xyzq)
    for i in range(5):
        a = xyzq.copy()
        a[i, 3] += 1e-3
        ep = gse_energy(orc, p, a)
        a[i, 3] -= 2e-3
        em = gse_energy(orc, p, a)
        assert abs((ep - em) / 2e-3 - r.pot[i]) < 1e-9 * np.max(np.abs(r.pot))


def test_virial_matches_strain_derivative(orc):
    """W_aa = -dE/d eps_aa with atoms and box scaled together (grid count
    fixed), by central differences; off-diagonal part symmetric."""
    L = 16.0
    p = orc.Params(16, L, 2.0)
    xyzq = safe_particles(8, L, 1.0, 23)
    r = orc.gse_step(p, xyzq)
    W = r.virial
    assert np.max(np.abs(W - W.T)) < 1e-10 * np.max(np.abs(W))
    eps = 1e-6
    for a in range(3):
        es = []
        for s in (1, -1):
            x = xyzq.copy()
            x[:, a] *= 1 + s * eps
            box = [L, L, L]
            box[a] *= 1 + s * eps
            q = orc.Params(16, tuple(box), 2.0)
            es.append(gse_energy(orc, q, x))
        fd = -(es[0] - es[1]) / (2 * eps)
        assert abs(fd - W[a, a]) < 1e-6 * np.max(np.abs(W)), (a, fd, W[a, a])


def test_energy_converges_to_ewald(orc, ref):
    """GSE-k approaches the Ewald reciprocal energy of the same sigma as the
    grid refines (cpu_gse.cpp:886-892 compares the two)."""
    xyzq = systems.pair(1.0, 20.0, orientation="diagonal")
    e_ew, f_ew = orc.ewald_recip(2.121320344, 32, (20.0,) * 3, xyzq)
    a = np.ascontiguousarray(xyzq)
    f = np.zeros((3, 2))
    e_ref = ref.ref_ewald_recip(ctypes.c_double(2.121320344), 32, *map(ctypes.c_double, (20.0,) * 3),
                                ctypes.c_double(332.0716), 2, P(a), P(f[0]), P(f[1]), P(f[2]))
    assert e_ew == e_ref and np.array_equal(f_ew, f.T)
    r = orc.gse_step(orc.Params(20, 20.0, 2.121320344), xyzq)
    assert abs(r.energy - e_ew) < 2e-3 * abs(e_ew)
    assert orc.erms(f_ew, r.force) < 5e-3


DATA_48K = "/root/reference/data/xyzq_48k.txt"


@pytest.mark.skipif(not os.path.exists(DATA_48K), reason="reference data not mounted")
def test_48k_energy_matches_stored_ewald_value(orc):
    """The reference stores the Ewald reciprocal energy of its 48,191-atom
    system, 1733.61 kcal/mol (cpu_gse.cpp:328); GSE-k at M=80 approximates it."""
    xyzq = np.loadtxt(DATA_48K)
    p = orc.Params(80, 80.0, 2.12132034355964)
    r = orc.gse_step(p, xyzq)
    assert abs(r.energy - 1733.61) < 1e-3 * 1733.61


def test_corrections_closed_form(orc):
    xyzq = np.array([[0.0, 0.0, 0.0, 1.0], [1.0, 2.0, 3.0, -3.0]])
    es, en = orc.corrections(2.0, (10.0, 11.0, 12.0), xyzq)
    assert abs(es - (-332.0716 * 10.0 / math.sqrt(2 * math.pi * 4.0))) < 1e-12
    assert abs(en - (-332.0716 * math.pi * 4.0 * 4.0 / 1320.0)) < 1e-12


def test_efield_and_dipole_vs_reference(orc, ref):
    dims, box = (10, 9, 8), (10.0, 9.9, 8.8)
    X, Y, Z = dims
    g = np.random.default_rng(2).normal(size=(Z, Y, X))
    mine = orc.efield(dims, box, g)
    e = np.zeros((3, Z, Y, X))
    ref.ref_efield_dd(X, Y, Z, *map(ctypes.c_double, box), P(g), P(e[0]), P(e[1]), P(e[2]))
    assert np.array_equal(mine, e)
    d = np.zeros(3)
    ref.ref_dipole_dd(X, Y, Z, *map(ctypes.c_double, box), P(g), P(d))
    assert np.array_equal(orc.dipole_sum(dims, box, g), d)


def test_openmp_timing_step_matches_serial_oracle(orc):
    """The OpenMP timing build of the step (the all-core CPU baseline) gives
    the serial oracle's energy, reference forces and phi_i to rounding, on
    one thread and on several."""
    from gse_amd import systems
    x = systems.water_box(400, 24.0, seed=12)
    p = orc.Params(24, 24.0, 2.12)
    r = orc.gse_step(p, x)
    for nt in (1, 3):
        orc.set_omp_threads(nt)
        e, f, pot, used = orc.step_omp(p, x)
        assert used == nt
        assert abs(e - r.energy) <= 1e-12 * abs(r.energy)
        assert np.max(np.abs(f - r.force)) <= 1e-12 * np.max(np.abs(r.force))
        assert np.max(np.abs(pot - r.pot)) <= 1e-12 * np.max(np.abs(r.pot))
