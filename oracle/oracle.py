"""numpy/ctypes binding of oracle/liboracle_gse.so (the FP64 CPU oracle).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Each wrapper names the reference routine that
the C function restates (see oracle/gse_oracle.c for file:line citations).

Particles are float64 arrays of shape (n, 4) with columns x, y, z, q.
Grids are float64 arrays of shape (Z, Y, X) (x fastest, as the reference).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None
_OMP = None

_dp = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int
_d = ctypes.c_double


def build(ref=False):
    """Compile the oracle (and with ref=True the reference library)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])
    if ref:
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_gse.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_support_size.restype = _i
        L.orc_energy.restype = _d
        L.orc_energy_kspace.restype = _d
        L.orc_ewald_recip.restype = _d
        _LIB = L
    return _LIB


def omp_lib():
    """OpenMP timing build of the step (oracle/liboracle_omp.so)."""
    global _OMP
    if _OMP is None:
        path = os.path.join(_HERE, "liboracle_omp.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_step_omp.restype = _d
        L.orc_omp_threads.restype = _i
        _OMP = L
    return _OMP


def set_omp_threads(n):
    omp_lib().orc_omp_set_threads(int(n))


def step_omp(params, xyzq):
    """One GSE-k step on all OpenMP threads (timing build; sums in another
    order than gse_step).  Returns (energy, force (n, 3) x ccelec, pot x
    ccelec, threads)."""
    a, n = _xyzq(xyzq)
    m = max(n, 1)
    f = np.zeros((3, m))
    pot = np.zeros(m)
    L = omp_lib()
    c = params._c()
    e = L.orc_step_omp(ctypes.byref(c), n, _p(a), _p(f[0]), _p(f[1]), _p(f[2]), _p(pot))
    return e, f[:, :n].T.copy(), pot[:n].copy(), L.orc_omp_threads()


def ref_lib():
    """The reference's own CPU classes (oracle/_ref), or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libgse_ref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_support_size.restype = _i
        L.ref_ewald_recip.restype = _d
        _REF = L
    return _REF


def _p(a):
    return a.ctypes.data_as(_dp)


def _xyzq(xyzq):
    a = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
    return a, a.shape[0]


def _grid(g):
    return np.ascontiguousarray(g, dtype=np.float64)


class Params:
    """CpuGSE constructor parameters (CpuGSE.h:52-55)."""

    def __init__(self, M, L, sigma, kappa=0.4, lambda_sigma=3.0,
                 lambda_sigma1=3.0, ccelec=332.0716):
        self.X, self.Y, self.Z = (M, M, M) if np.isscalar(M) else tuple(M)
        self.bx, self.by, self.bz = (L, L, L) if np.isscalar(L) else tuple(L)
        self.sigma = float(sigma)
        self.kappa = float(kappa)
        self.lambda_sigma = float(lambda_sigma)
        self.lambda_sigma1 = float(lambda_sigma1)
        self.ccelec = float(ccelec)

    # CpuGSE.cpp:33-44 and the rcut of CpuGSEk.cpp:31 / CpuGSEr.cpp:40
    @property
    def sigma1(self):
        return self.sigma * np.sqrt(self.kappa)

    @property
    def sigma2(self):
        return self.sigma * np.sqrt(1.0 - 2.0 * self.kappa)

    @property
    def rcut1(self):
        return self.lambda_sigma1 * np.sqrt(2.0) * self.sigma1

    @property
    def rcut2(self):
        return self.lambda_sigma1 * np.sqrt(2.0) * self.sigma2

    @property
    def dims(self):
        return self.X, self.Y, self.Z

    @property
    def box(self):
        return self.bx, self.by, self.bz

    def _c(self):
        class S(ctypes.Structure):
            _fields_ = [("X", _i), ("Y", _i), ("Z", _i), ("sigma", _d),
                        ("kappa", _d), ("lambda_sigma", _d),
                        ("lambda_sigma1", _d), ("bx", _d), ("by", _d),
                        ("bz", _d), ("ccelec", _d)]
        return S(self.X, self.Y, self.Z, self.sigma, self.kappa,
                 self.lambda_sigma, self.lambda_sigma1, self.bx, self.by,
                 self.bz, self.ccelec)


def support_size(dims, rcut, box):
    return lib().orc_support_size(*dims, _d(rcut), *map(_d, box))


def spread(dims, sigma, rcut, xyzq, box, normalize=False):
    """CpuGaussCharge::spreadChargeToGrid (CpuGaussCharge.cpp:235-384)."""
    a, n = _xyzq(xyzq)
    X, Y, Z = dims
    rho = np.zeros((Z, Y, X))
    lib().orc_spread(X, Y, Z, _d(sigma), _d(rcut), n, _p(a), *map(_d, box),
                     _p(rho), int(bool(normalize)))
    return rho


def spread_on_grid(dims, sigma, rcut, box, rho):
    """CpuGaussCharge::spreadChargeOnGrid (CpuGaussCharge.cpp:390-519)."""
    X, Y, Z = dims
    r = _grid(rho)
    out = np.zeros((Z, Y, X))
    lib().orc_spread_on_grid(X, Y, Z, _d(sigma), _d(rcut), *map(_d, box),
                             _p(r), _p(out))
    return out


def conv_run_x(dims, kx, grid):
    """CpuConvolution::run (CpuConvolution.cpp:116-151), X pass only."""
    X, Y, Z = dims
    k = np.ascontiguousarray(kx, dtype=np.float64)
    g = _grid(grid)
    out = np.zeros((Z, Y, X))
    lib().orc_conv_run_x(X, Y, Z, (k.size - 1) // 2, _p(k), _p(g), _p(out))
    return out


def conv_run_xyz(dims, kx, ky, kz, grid):
    X, Y, Z = dims
    ks = [np.ascontiguousarray(k, dtype=np.float64) for k in (kx, ky, kz)]
    g = _grid(grid)
    out = np.zeros((Z, Y, X))
    lib().orc_conv_run_xyz(X, Y, Z, *[(k.size - 1) // 2 for k in ks],
                           *[_p(k) for k in ks], _p(g), _p(out))
    return out


def interpolate(dims, sigma, rcut, xyzq, box, phi):
    """CpuGaussCharge::interpolateLoop (CpuGaussCharge.cpp:105-172)."""
    a, n = _xyzq(xyzq)
    X, Y, Z = dims
    g = _grid(phi)
    f = np.zeros((3, max(n, 1)))
    lib().orc_interpolate(X, Y, Z, _d(sigma), _d(rcut), n, _p(a),
                          *map(_d, box), _p(g), _p(f[0]), _p(f[1]), _p(f[2]))
    return f[:, :n].T.copy()


def gather_cube(dims, sigma, rcut, xyzq, box, phi):
    """Cube-support gather: (pot[n], force[n,3], vir[3,3]), no ccelec."""
    a, n = _xyzq(xyzq)
    X, Y, Z = dims
    g = _grid(phi)
    pot = np.zeros(max(n, 1))
    f = np.zeros((3, max(n, 1)))
    v = np.zeros(9)
    lib().orc_gather_cube(X, Y, Z, _d(sigma), _d(rcut), n, _p(a),
                          *map(_d, box), _p(g), _p(pot), _p(f[0]), _p(f[1]),
                          _p(f[2]), _p(v))
    return pot[:n].copy(), f[:, :n].T.copy(), v.reshape(3, 3)


def fft_r2c(grid):
    g = _grid(grid)
    Z, Y, X = g.shape
    out = np.zeros((Z, Y, X // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(X, Y, Z, _p(g), out.ctypes.data_as(_dp))
    return out


def fft_c2r(spec, X):
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    Z, Y, _ = s.shape
    out = np.zeros((Z, Y, X))
    lib().orc_fft_c2r(X, Y, Z, s.ctypes.data_as(_dp), _p(out))
    return out


def fft_1d(data, sign):
    d = np.ascontiguousarray(data, dtype=np.complex128).copy()
    lib().orc_fft_1d(d.size, d.ctypes.data_as(_dp), int(sign))
    return d


def greens_mul(kind, dims, box, sigma, spec):
    """CpuGreensFuncGSE::mul (kind 0) / CpuGreensFuncG2::mul (kind 1)."""
    s = np.ascontiguousarray(spec, dtype=np.complex128).copy()
    lib().orc_greens_mul(kind, *dims, *map(_d, box), _d(sigma),
                         s.ctypes.data_as(_dp))
    return s


def fft_solve(kind, dims, sigma, box, rho):
    """CpuFFTSolver::run (CpuFFTSolver.cpp:75-95)."""
    X, Y, Z = dims
    r = _grid(rho)
    phi = np.zeros((Z, Y, X))
    lib().orc_fft_solve(kind, X, Y, Z, _d(sigma), *map(_d, box), _p(r),
                        _p(phi))
    return phi


def energy(dims, ccelec, box, rho, phi):
    """CpuGSE::calculateEnergy (CpuGSE.cpp:86-104)."""
    return lib().orc_energy(*dims, _d(ccelec), *map(_d, box),
                            _p(_grid(rho)), _p(_grid(phi)))


def energy_kspace(kind, dims, ccelec, box, sigma, spec):
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    v = np.zeros(9)
    e = lib().orc_energy_kspace(kind, *dims, _d(ccelec), *map(_d, box),
                                _d(sigma), s.ctypes.data_as(_dp), _p(v))
    return e, v.reshape(3, 3)


def efield(dims, box, phi):
    X, Y, Z = dims
    out = np.zeros((3, Z, Y, X))
    lib().orc_efield(X, Y, Z, *map(_d, box), _p(_grid(phi)), _p(out[0]),
                     _p(out[1]), _p(out[2]))
    return out


def dipole_sum(dims, box, rho):
    d = np.zeros(3)
    lib().orc_dipole_sum(*dims, *map(_d, box), _p(_grid(rho)), _p(d))
    return d


def ewald_recip(sigma, kcut, box, xyzq, ccelec=332.0716):
    """CpuEwaldRecip::calcForceEnergy (CpuEwaldRecip.cpp:87-147)."""
    a, n = _xyzq(xyzq)
    f = np.zeros((3, max(n, 1)))
    e = lib().orc_ewald_recip(_d(sigma), int(kcut), *map(_d, box),
                              _d(ccelec), n, _p(a), _p(f[0]), _p(f[1]),
                              _p(f[2]))
    return e, f[:, :n].T.copy()


def corrections(sigma, box, xyzq, ccelec=332.0716):
    a, n = _xyzq(xyzq)
    es, en = _d(), _d()
    lib().orc_corrections(_d(sigma), *map(_d, box), _d(ccelec), n, _p(a),
                          ctypes.byref(es), ctypes.byref(en))
    return es.value, en.value


class StepResult:
    pass


def gse_step(params, xyzq, mode=0, want_grids=False):
    """Full GSE step (CpuGSEk spread/solve/energy/force, cpu_gse.cpp:719-735).

    Returns an object with energy, energy_k, force (sphere, ccelec
    included), force_cube, pot, virial and optionally rho, phi.
    """
    a, n = _xyzq(xyzq)
    X, Y, Z = params.dims
    m = max(n, 1)
    e, ek = _d(), _d()
    f = np.zeros((3, m))
    fc = np.zeros((3, m))
    pot = np.zeros(m)
    vir = np.zeros(9)
    rho = np.zeros((Z, Y, X)) if want_grids else None
    phi = np.zeros((Z, Y, X)) if want_grids else None
    c = params._c()
    lib().orc_gse_step(ctypes.byref(c), int(mode), n, _p(a), ctypes.byref(e),
                       ctypes.byref(ek), _p(f[0]), _p(f[1]), _p(f[2]),
                       _p(fc[0]), _p(fc[1]), _p(fc[2]), _p(pot), _p(vir),
                       _p(rho) if want_grids else None,
                       _p(phi) if want_grids else None)
    r = StepResult()
    r.energy = e.value
    r.energy_k = ek.value
    r.force = f[:, :n].T.copy()
    r.force_cube = fc[:, :n].T.copy()
    r.pot = pot[:n].copy()
    r.virial = vir.reshape(3, 3)
    r.rho = rho
    r.phi = phi
    return r


def erms(ref, x):
    """Relative force RMS error, cpu_gse.cpp:131-146 (Erms)."""
    ref = np.asarray(ref, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    return float(np.sqrt(np.sum((x - ref) ** 2) / np.sum(ref ** 2)))


# ---------------------------------------------------------------- Ewald direct
def _excl(iblo14, inb14):
    if iblo14 is None:
        return 0, None, None
    ib = np.ascontiguousarray(iblo14, dtype=np.int32)
    nb = np.ascontiguousarray(inb14 if len(inb14) else [0], dtype=np.int32)
    return len(ib), ib, nb


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def ewald_direct(sigma, rcut, box, xyzq, iblo14=None, inb14=None, ccelec=332.0716,
                 reference=False):
    """CpuEwaldDirect::calcForceEnergy (CpuEwaldDirect.cpp:45-131): returns
    (E_direct + E_self, dE/dr as an (n, 3) array).  reference=True runs the
    reference's own compiled class (oracle/_ref) instead of the restatement."""
    a, n = _xyzq(xyzq)
    ne, ib, nb = _excl(iblo14, inb14)
    g = np.zeros((3, max(n, 1)))
    if reference:
        L = ref_lib()
        fn = L.ref_ewald_direct
    else:
        L = lib()
        fn = L.orc_ewald_direct
    fn.restype = _d
    e = fn(_d(sigma), _d(rcut), *map(_d, box), _d(ccelec), n, _p(a), ne, _ip(ib), _ip(nb),
           _p(g[0]), _p(g[1]), _p(g[2]))
    return e, g[:, :n].T.copy()


# ------------------------------------------------------------------- SPME
def pme_theta(order, w):
    th = np.zeros(16)
    dth = np.zeros(16)
    lib().orc_pme_theta(int(order), _d(w), _p(th), _p(dth))
    return th[:order].copy(), dth[:order].copy()


def pme_prefac(order, nfft):
    p = np.zeros(nfft)
    lib().orc_pme_prefac(int(order), int(nfft), _p(p))
    return p


def pme_spread(dims, order, xyzq, box):
    a, n = _xyzq(xyzq)
    X, Y, Z = dims
    Q = np.zeros((Z, Y, X))
    lib().orc_pme_spread(X, Y, Z, int(order), n, _p(a), *map(_d, box), _p(Q))
    return Q


def pme_step(dims, order, xyzq, box, kappa, ccelec=332.0716, want_grid=False):
    """SPME reciprocal step as gpu_gse.cu:61-73 drives CudaPMERecip: returns
    (energy, virial 3x3, forces (n, 3) = -dE/dr[, the charge grid])."""
    a, n = _xyzq(xyzq)
    X, Y, Z = dims
    e = _d()
    vir = np.zeros(9)
    f = np.zeros((3, max(n, 1)))
    Q = np.zeros((Z, Y, X)) if want_grid else None
    lib().orc_pme_step(X, Y, Z, int(order), n, _p(a), *map(_d, box), _d(kappa), _d(ccelec),
                       ctypes.byref(e), _p(vir), _p(f[0]), _p(f[1]), _p(f[2]),
                       _p(Q) if want_grid else None)
    out = (e.value, vir.reshape(3, 3), f[:, :n].T.copy())
    return out + (Q,) if want_grid else out


# ------------------------------------------------------------ GSE-u, GSE-r
def gser_step(params, xyzq, nsweep=0):
    """GSE-r (CpuGSEr.cpp): nsweep > 0 relaxes that many red-black sweeps
    from phi = 0 as the reference does, nsweep = 0 solves exactly.  Returns
    (energy, forces (n, 3) x ccelec, rho, rhoS, phi)."""
    a, n = _xyzq(xyzq)
    X, Y, Z = params.dims
    e = _d()
    f = np.zeros((3, max(n, 1)))
    g = [np.zeros((Z, Y, X)) for _ in range(3)]
    c = params._c()
    lib().orc_gser_step(ctypes.byref(c), int(nsweep), n, _p(a), ctypes.byref(e), _p(f[0]),
                        _p(f[1]), _p(f[2]), _p(g[0]), _p(g[1]), _p(g[2]))
    return e.value, f[:, :n].T.copy(), g[0], g[1], g[2]


def gseu_step(params, N, b, xyzq):
    """GSE-u (CpuGSEu.cpp): returns (energy, forces (n, 3), phi)."""
    a, n = _xyzq(xyzq)
    X, Y, Z = params.dims
    e = _d()
    f = np.zeros((3, max(n, 1)))
    phi = np.zeros((Z, Y, X))
    c = params._c()
    lib().orc_gseu_step(int(N), _d(b), ctypes.byref(c), n, _p(a), ctypes.byref(e), _p(f[0]),
                        _p(f[1]), _p(f[2]), _p(phi))
    return e.value, f[:, :n].T.copy(), phi


def mehrstellen_solve(dims, box, rhoS):
    X, Y, Z = dims
    phi = np.zeros((Z, Y, X))
    lib().orc_mehrstellen_solve(X, Y, Z, *map(_d, box), _p(_grid(rhoS)), _p(phi))
    return phi


def rbgs_hermitian(dims, box, rho, phi0, nsweep):
    X, Y, Z = dims
    phi = np.ascontiguousarray(phi0, dtype=np.float64).copy()
    lib().orc_rbgs_hermitian(X, Y, Z, *map(_d, box), _p(_grid(rho)), _p(phi), int(nsweep))
    return phi


def ref_gseu_step(params, N, b, xyzq):
    """The reference's own CpuGSEu (oracle/_ref); writes rhoG%d.txt into the
    working directory as the reference does."""
    a, n = _xyzq(xyzq)
    X, Y, Z = params.dims
    f = np.zeros((3, max(n, 1)))
    phi = np.zeros((Z, Y, X))
    L = ref_lib()
    L.ref_gseu_step.restype = _d
    e = L.ref_gseu_step(int(N), _d(b), X, Y, Z, _d(params.sigma), _d(params.kappa),
                        _d(params.lambda_sigma), _d(params.lambda_sigma1), *map(_d, params.box),
                        n, _p(a), _p(f[0]), _p(f[1]), _p(f[2]), _p(phi))
    return e, f[:, :n].T.copy(), phi


def ref_gser_step(params, xyzq):
    """The reference's own CpuGSEr with its 200-sweep solver (oracle/_ref);
    writes phiCoarsest.txt, rhoCoarsest.txt and resCoarsest.txt into the
    working directory.  Returns (energy, forces, rhoS, phi)."""
    a, n = _xyzq(xyzq)
    X, Y, Z = params.dims
    f = np.zeros((3, max(n, 1)))
    rhoS = np.zeros((Z, Y, X))
    phi = np.zeros((Z, Y, X))
    L = ref_lib()
    L.ref_gser_step.restype = _d
    e = L.ref_gser_step(X, Y, Z, _d(params.sigma), _d(params.kappa), _d(params.lambda_sigma),
                        _d(params.lambda_sigma1), *map(_d, params.box), n, _p(a), _p(f[0]),
                        _p(f[1]), _p(f[2]), _p(rhoS), _p(phi))
    return e, f[:, :n].T.copy(), rhoS, phi
