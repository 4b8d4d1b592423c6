"""Python mirror of the reference's class API for the GSE k-space path, on
device-resident torch tensors, backed by the HIP library (gse_amd/_lib.py).

Class and method names follow the reference (CpuGaussCharge.h,
CpuConvolution.h, CpuFFTSolver.h, CpuGreensFunc.h, CpuGSE.h, CpuGSEk.h) so
parity tests read like the reference's own drivers; every method runs on the
GPU through include/gse_hip.h.  Tensors:
  xyzq      (N, 4) float32/float64 {x, y, z, q}, contiguous, on the device
  grids     (Z, Y, X) of the compute type (x fastest)
  forces    (N,) arrays of the accumulation type
All calls run on torch's current stream.
"""
import ctypes

import torch

from . import _lib
from ._lib import GSE_FLOAT32, GSE_FLOAT64, GREENS_G2, GREENS_GSE, GREENS_HERMITIAN, check  # noqa: F401

_DT = {torch.float32: GSE_FLOAT32, torch.float64: GSE_FLOAT64}
_TD = {GSE_FLOAT32: torch.float32, GSE_FLOAT64: torch.float64}


def _code(dtype):
    if isinstance(dtype, int):
        return dtype
    return _DT[dtype]


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must be on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _box(box):
    if isinstance(box, (int, float)):
        return (float(box),) * 3
    return tuple(float(b) for b in box)


def _dims(M):
    if isinstance(M, int):
        return (M, M, M)
    return tuple(int(m) for m in M)


def _check_xyzq(xyzq, ct):
    if xyzq.dim() != 2 or xyzq.shape[1] != 4:
        raise ValueError("xyzq must have shape (N, 4)")
    if xyzq.dtype != ct:
        raise TypeError(f"xyzq must be {ct}, got {xyzq.dtype}")


def _check_count(numCoord, xyzq):
    """numCoord may not exceed the records xyzq holds: the kernels read
    numCoord records from the device pointer."""
    if not isinstance(numCoord, int) or numCoord < 0 or numCoord > xyzq.shape[0]:
        raise ValueError(f"numCoord must be in [0, {xyzq.shape[0]}] (got {numCoord})")


def _layout(fn, h):
    info = (ctypes.c_int * _lib.LAYOUT_FIELDS)()
    check(fn(h, info))
    v = list(info)
    return {"bin": tuple(v[0:3]), "nbins": v[3], "brick": tuple(v[4:7]), "direct": v[7],
            "reduce": _lib.REDUCE_PATHS[v[8]], "cap_bin": v[9],
            "kspace": _lib.KSPACE_FORMS[v[10]],
            "spread": {-1: "none", 0: "lds_atomic", 1: "mma"}[v[11]]}


def _check_grid(g, dims, dt):
    X, Y, Z = dims
    if tuple(g.shape) != (Z, Y, X):
        raise ValueError(f"grid must have shape {(Z, Y, X)}, got {tuple(g.shape)}")
    if g.dtype != dt:
        raise TypeError(f"grid must be {dt}, got {g.dtype}")


class CpuGaussCharge:
    """CpuGaussCharge<AT,CT> (CpuGaussCharge.h:11-49) on the GPU."""

    def __init__(self, at, ct, sizeX, sizeY=None, sizeZ=None):
        L = _lib.load()
        self.at = _TD[_code(at)]
        self.ct = _TD[_code(ct)]
        self.dims = (sizeX, sizeY or sizeX, sizeZ or sizeX)
        h = ctypes.c_void_p()
        check(L.gse_gauss_charge_create(ctypes.byref(h), _code(at), _code(ct), *self.dims))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_gauss_charge_destroy(self._h)
            self._h = None

    def _s(self):
        check(self._L.gse_gauss_charge_set_stream(self._h, _stream()))

    def calcSupportSize(self, rcut, boxx, boxy, boxz):
        n = ctypes.c_int()
        check(self._L.gse_gauss_charge_calc_support_size(self._h, rcut, boxx, boxy, boxz,
                                                         ctypes.byref(n)))
        return n.value

    def spreadChargeToGrid(self, sigma, rcut, numCoord, xyzq, boxx, boxy, boxz, rho,
                           normalize=False):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        _check_grid(rho, self.dims, self.ct)
        self._s()
        check(self._L.gse_gauss_charge_spread_charge_to_grid(
            self._h, sigma, rcut, numCoord, _ptr(xyzq), boxx, boxy, boxz, _ptr(rho),
            int(bool(normalize))))

    def spreadChargeOnGrid(self, sigma, rcut, boxx, boxy, boxz, rho, rhoS):
        _check_grid(rho, self.dims, self.ct)
        _check_grid(rhoS, self.dims, self.ct)
        self._s()
        check(self._L.gse_gauss_charge_spread_charge_on_grid(self._h, sigma, rcut, boxx, boxy,
                                                             boxz, _ptr(rho), _ptr(rhoS)))

    def interpolateForce(self, sigma, rcut, numCoord, xyzq, boxx, boxy, boxz, phi, forceX,
                         forceY, forceZ):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        _check_grid(phi, self.dims, self.ct)
        for f in (forceX, forceY, forceZ):
            if f.dtype != self.at or f.numel() < numCoord:
                raise ValueError("force arrays must be AT with numCoord entries")
        self._s()
        check(self._L.gse_gauss_charge_interpolate_force(
            self._h, sigma, rcut, numCoord, _ptr(xyzq), boxx, boxy, boxz, _ptr(phi),
            _ptr(forceX), _ptr(forceY), _ptr(forceZ)))

    def calcElectricFieldOnGrid(self, boxx, boxy, boxz, phi, Ex, Ey, Ez):
        for g in (phi, Ex, Ey, Ez):
            _check_grid(g, self.dims, self.ct)
        self._s()
        check(self._L.gse_gauss_charge_calc_electric_field_on_grid(
            self._h, boxx, boxy, boxz, _ptr(phi), _ptr(Ex), _ptr(Ey), _ptr(Ez)))

    def calcDipoleSum(self, boxx, boxy, boxz, rho):
        _check_grid(rho, self.dims, self.ct)
        self._s()
        out = (ctypes.c_double * 3)()
        check(self._L.gse_gauss_charge_calc_dipole_sum(self._h, boxx, boxy, boxz, _ptr(rho),
                                                       out))
        return tuple(out)

    def gather(self, sigma, rcut, xyzq, box, phi, what, scale=1.0, out=None):
        """Build-defined gather: returns dict of the requested outputs."""
        _check_xyzq(xyzq, self.ct)
        n = xyzq.shape[0]
        dev = xyzq.device
        o = {} if out is None else dict(out)
        d = _lib.GatherDesc()
        d.what = what
        d.scale = scale
        if what & _lib.GATHER_SPHERE_FORCE:
            o.setdefault("force", torch.empty((3, n), dtype=self.at, device=dev))
            d.forceX, d.forceY, d.forceZ = (_ptr(o["force"][k]) for k in range(3))
        if what & _lib.GATHER_CUBE_FORCE:
            o.setdefault("force_cube", torch.empty((3, n), dtype=self.at, device=dev))
            d.cubeX, d.cubeY, d.cubeZ = (_ptr(o["force_cube"][k]) for k in range(3))
        if what & _lib.GATHER_POTENTIAL:
            o.setdefault("pot", torch.empty(n, dtype=self.at, device=dev))
            d.potential = _ptr(o["pot"])
        if what & _lib.GATHER_VIRIAL:
            o.setdefault("virial", torch.empty(9, dtype=torch.float64, device=dev))
            d.virial = _ptr(o["virial"])
        self._s()
        check(self._L.gse_gauss_charge_gather(self._h, sigma, rcut, n, _ptr(xyzq), *_box(box),
                                              _ptr(phi), ctypes.byref(d)))
        return o

    def invalidCount(self):
        c = ctypes.c_int()
        self._s()
        check(self._L.gse_gauss_charge_invalid_count(self._h, ctypes.byref(c)))
        return c.value

    def layout(self):
        """Bin layout of the last spread or gather and the kernel that
        converted the last spread's fixed-point sums (host only)."""
        return _layout(self._L.gse_gauss_charge_get_layout, self._h)


class CpuConvolution:
    """CpuConvolution<T> (CpuConvolution.h:9-38) on the GPU."""

    def __init__(self, t, sizeX, sizeY=None, sizeZ=None):
        L = _lib.load()
        self.t = _TD[_code(t)]
        self.dims = (sizeX, sizeY or sizeX, sizeZ or sizeX)
        h = ctypes.c_void_p()
        check(L.gse_convolution_create(ctypes.byref(h), _code(t), *self.dims))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_convolution_destroy(self._h)
            self._h = None

    def setupKernel(self, kernelRadiusX, kernelRadiusY, kernelRadiusZ, h_kernelX, h_kernelY,
                    h_kernelZ):
        import numpy as np
        npdt = np.float64 if self.t == torch.float64 else np.float32
        ks = [np.ascontiguousarray(k, dtype=npdt) for k in (h_kernelX, h_kernelY, h_kernelZ)]
        for k, r in zip(ks, (kernelRadiusX, kernelRadiusY, kernelRadiusZ)):
            if k.size < 2 * r + 1:
                raise ValueError("kernel arrays need 2R+1 taps")
        self._keep = ks
        check(self._L.gse_convolution_setup_kernel(
            self._h, kernelRadiusX, kernelRadiusY, kernelRadiusZ,
            *[ctypes.c_void_p(k.ctypes.data) for k in ks]))

    def run(self, data, dataXYZ):
        _check_grid(data, self.dims, self.t)
        _check_grid(dataXYZ, self.dims, self.t)
        check(self._L.gse_convolution_set_stream(self._h, _stream()))
        check(self._L.gse_convolution_run(self._h, _ptr(data), _ptr(dataXYZ)))

    def runXYZ(self, data, dataXYZ):
        _check_grid(data, self.dims, self.t)
        _check_grid(dataXYZ, self.dims, self.t)
        check(self._L.gse_convolution_set_stream(self._h, _stream()))
        check(self._L.gse_convolution_run_xyz(self._h, _ptr(data), _ptr(dataXYZ)))


class CpuGreensFuncGSE:
    """CpuGreensFuncGSE<T> (CpuGreensFuncGSE.h): kind tag for CpuFFTSolver,
    and mul() on a device half spectrum (CpuGreensFunc.h:16-18)."""
    kind = GREENS_GSE

    def mul(self, nfftX, nfftY, nfftZ, boxx, boxy, boxz, sigma, rho):
        t = GSE_FLOAT64 if rho.dtype in (torch.complex128, torch.float64) else GSE_FLOAT32
        check(_lib.load().gse_greens_func_mul(t, self.kind, nfftX, nfftY, nfftZ, boxx, boxy,
                                              boxz, sigma, _ptr(rho), _stream()))


class CpuGreensFuncG2(CpuGreensFuncGSE):
    """CpuGreensFuncG2<T> (CpuGreensFuncG2.h)."""
    kind = GREENS_G2


class CpuGreensFuncHermitian(CpuGreensFuncGSE):
    """The HERMITIAN (Mehrstellen) finite-difference Poisson operator of
    CpuMultiGridSolver.cpp:237-251 as a Green's function: phi = A^-1 B
    (-4 pi rho), zero mean; sigma is not used."""
    kind = GREENS_HERMITIAN


class CpuFFTSolver:
    """CpuFFTSolver<T> (CpuFFTSolver.h:9-43) on the GPU with rocFFT."""

    def __init__(self, greensFunc, t, sizeX, sizeY=None, sizeZ=None):
        L = _lib.load()
        self.t = _TD[_code(t)]
        self.dims = (sizeX, sizeY or sizeX, sizeZ or sizeX)
        h = ctypes.c_void_p()
        check(L.gse_fft_solver_create(ctypes.byref(h), _code(t), greensFunc.kind, *self.dims))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_fft_solver_destroy(self._h)
            self._h = None

    def run(self, phi, rho, sigma, boxx, boxy, boxz):
        _check_grid(phi, self.dims, self.t)
        _check_grid(rho, self.dims, self.t)
        check(self._L.gse_fft_solver_set_stream(self._h, _stream()))
        check(self._L.gse_fft_solver_run(self._h, _ptr(phi), _ptr(rho), sigma, boxx, boxy, boxz))

    def r2c(self, rho, rhoK):
        check(self._L.gse_fft_solver_set_stream(self._h, _stream()))
        check(self._L.gse_fft_solver_r2c(self._h, _ptr(rho), _ptr(rhoK)))

    def c2r(self, rhoK, phi):
        check(self._L.gse_fft_solver_set_stream(self._h, _stream()))
        check(self._L.gse_fft_solver_c2r(self._h, _ptr(rhoK), _ptr(phi)))

    def form(self):
        """"fused": run() takes 2D rocFFT transforms around the fused z
        pass; "fused4": the same with the four-step register z pass;
        "fused/xy", "fused4/xy": those z passes between the plane kernels'
        x-y transforms; "fused/y4", "fused4/y4": between 1D x rocFFT
        transforms and the four-step y pass; "3d": 3D transforms and a
        separate k-pass."""
        v = ctypes.c_int()
        check(self._L.gse_fft_solver_get_form(self._h, ctypes.byref(v)))
        return _lib.KSPACE_FORMS[v.value]


class CpuGSEk:
    """CpuGSEk<AT,CT> (CpuGSEk.h:18-37, base CpuGSE.h:18-68) on the GPU, plus
    the build-defined fused step, adjoint and corrections."""

    def __init__(self, at, ct, sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma, lambdaSigma1,
                 boxx, boxy, boxz, ccelec=332.0716):
        L = _lib.load()
        self.at = _TD[_code(at)]
        self.ct = _TD[_code(ct)]
        self.dims = (sizeX, sizeY, sizeZ)
        h = ctypes.c_void_p()
        check(L.gse_gsek_create(ctypes.byref(h), _code(at), _code(ct), sizeX, sizeY, sizeZ,
                                sigma, kappa, lambdaSigma, lambdaSigma1, boxx, boxy, boxz,
                                ccelec))
        self._h = h
        self._L = L
        self.ccelec = ccelec

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_gsek_destroy(self._h)
            self._h = None

    def _s(self):
        check(self._L.gse_gsek_set_stream(self._h, _stream()))

    def setBoxSize(self, boxx, boxy, boxz):
        check(self._L.gse_gsek_set_box_size(self._h, boxx, boxy, boxz))

    def getSigma1(self):
        v = ctypes.c_double()
        check(self._L.gse_gsek_get_params(self._h, ctypes.byref(v), None, None))
        return v.value

    def getSigma2(self):
        v = ctypes.c_double()
        check(self._L.gse_gsek_get_params(self._h, None, ctypes.byref(v), None))
        return v.value

    def getLambdaSigma1(self):
        v = ctypes.c_double()
        check(self._L.gse_gsek_get_params(self._h, None, None, ctypes.byref(v)))
        return v.value

    def spreadCharge1(self, numCoord, xyzq):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        self._s()
        check(self._L.gse_gsek_spread_charge1(self._h, numCoord, _ptr(xyzq)))

    def spreadCharge2(self, rhoOut):
        _check_grid(rhoOut, self.dims, self.ct)
        self._s()
        check(self._L.gse_gsek_spread_charge2(self._h, _ptr(rhoOut)))

    def solvePoisson(self):
        self._s()
        check(self._L.gse_gsek_solve_poisson(self._h))

    def calculateEnergy(self):
        e = ctypes.c_double()
        self._s()
        check(self._L.gse_gsek_calculate_energy(self._h, ctypes.byref(e)))
        return e.value

    def interpolateForce(self, numCoord, xyzq, forceX, forceY, forceZ):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        for f in (forceX, forceY, forceZ):
            if f.dtype != self.at or f.numel() < numCoord:
                raise ValueError("force arrays must be AT with numCoord entries")
        self._s()
        check(self._L.gse_gsek_interpolate_force(self._h, numCoord, _ptr(xyzq), _ptr(forceX),
                                                 _ptr(forceY), _ptr(forceZ)))

    def _copy_grid(self, which):
        X, Y, Z = self.dims
        out = torch.empty((Z, Y, X), dtype=self.ct, device="cuda")
        self._s()
        check(self._L.gse_gsek_copy_grid(self._h, which, _ptr(out)))
        return out

    def layout(self):
        """Bin layout and conversion kernel of the handle's last spread."""
        return _layout(self._L.gse_gsek_get_layout, self._h)

    def getRhoS(self):
        """Copy of rhoS (CpuGSE.h:63)."""
        return self._copy_grid(0)

    def getPhi(self):
        """Copy of phiM (CpuGSE.h:64)."""
        return self._copy_grid(1)

    def step(self, xyzq, mode=0, force=True, potential=True, cube=False, virial=False,
             out=None):
        """Fused GSE step; returns dict with energy (device float64 scalar
        tensor) and the requested per-particle arrays."""
        _check_xyzq(xyzq, self.ct)
        n = xyzq.shape[0]
        dev = xyzq.device
        o = {} if out is None else out
        d = _lib.StepDesc()
        d.mode = mode
        if force:
            if "force" not in o:
                o["force"] = torch.empty((3, n), dtype=self.at, device=dev)
            d.forceX, d.forceY, d.forceZ = (_ptr(o["force"][k]) for k in range(3))
        if potential:
            if "pot" not in o:
                o["pot"] = torch.empty(n, dtype=self.at, device=dev)
            d.potential = _ptr(o["pot"])
        if cube:
            if "force_cube" not in o:
                o["force_cube"] = torch.empty((3, n), dtype=self.at, device=dev)
            d.cubeX, d.cubeY, d.cubeZ = (_ptr(o["force_cube"][k]) for k in range(3))
        if "energy" not in o:
            o["energy"] = torch.empty(1, dtype=torch.float64, device=dev)
        d.energy = _ptr(o["energy"])
        if virial:
            if "virial" not in o:
                o["virial"] = torch.empty(9, dtype=torch.float64, device=dev)
            d.virial = _ptr(o["virial"])
        self._s()
        check(self._L.gse_gsek_step(self._h, n, _ptr(xyzq), ctypes.byref(d)))
        return o

    def adjoint(self, xyzq, grad_energy=None, want_q=True, want_r=True, want_virial=True):
        """Backward sweep: dE/dq, dE/dr (scaled by grad_energy) and W."""
        _check_xyzq(xyzq, self.ct)
        n = xyzq.shape[0]
        dev = xyzq.device
        o = {}
        if want_q:
            o["dEdq"] = torch.empty(n, dtype=self.at, device=dev)
        if want_r:
            o["dEdr"] = torch.empty((3, n), dtype=self.at, device=dev)
        if want_virial:
            o["virial"] = torch.empty(9, dtype=torch.float64, device=dev)
        g = None
        if grad_energy is not None:
            g = grad_energy.to(device=dev, dtype=torch.float64).reshape(1).contiguous()
        self._s()
        r = o.get("dEdr")
        check(self._L.gse_gsek_adjoint(
            self._h, n, _ptr(xyzq), _ptr(g), _ptr(o.get("dEdq")),
            _ptr(r[0]) if r is not None else None, _ptr(r[1]) if r is not None else None,
            _ptr(r[2]) if r is not None else None, _ptr(o.get("virial"))))
        return o

    STAGES = ("bin", "spread", "convert", "r2c", "kpass", "c2r", "gather")

    def setProfiling(self, enable):
        check(self._L.gse_gsek_set_profiling(self._h, int(bool(enable))))

    def stageTimes(self):
        """Mean ms per stage over the steps recorded since profiling was
        enabled (or last read); synchronises."""
        ms = (ctypes.c_double * len(self.STAGES))()
        n = ctypes.c_int()
        check(self._L.gse_gsek_stage_times(self._h, ms, ctypes.byref(n)))
        return dict(zip(self.STAGES, list(ms))), n.value

    def corrections(self, xyzq):
        _check_xyzq(xyzq, self.ct)
        out = torch.empty(2, dtype=torch.float64, device=xyzq.device)
        self._s()
        check(self._L.gse_gsek_corrections(self._h, xyzq.shape[0], _ptr(xyzq), _ptr(out)))
        return out


def set_tiling(bin_edge=0, lds_budget=0):
    check(_lib.load().gse_set_tiling(int(bin_edge), int(lds_budget)))


TUNING_KEYS = ("bin_edge", "brick_budget", "spread_block", "gather_block", "spread_stage",
               "gather_stage", "spread_lds32", "spread_cols", "spread_flush",
               "reduce_tile", "reduce_nt", "reduce_tile2", "bin_agg", "fused_kz", "spread_mma",
               "zpass_cols", "gather_group", "zpass4", "xy_plane",
               "zpass4_cols", "bin_coarse_block", "ypass4",
               "bin_order")


def set_tuning(**kw):
    """Set named launch-shape knobs (see include/gse_hip.h)."""
    L = _lib.load()
    for k, v in kw.items():
        check(L.gse_set_tuning_param(k.encode(), int(v)))


def get_tuning():
    L = _lib.load()
    out = {}
    for k in TUNING_KEYS:
        v = ctypes.c_int()
        check(L.gse_get_tuning_param(k.encode(), ctypes.byref(v)))
        out[k] = v.value
    return out


class GSEBatch:
    """Many independent GSE-k systems stepped together (include/gse_hip.h,
    gse_batch_*): one launch per stage for all systems, optionally replayed
    from a HIP graph.  systems: list of (M or (X, Y, Z), L or (bx, by, bz),
    n_particles); particles of all systems are concatenated in list order."""

    def __init__(self, at, ct, systems, sigma, kappa=0.4, lambdaSigma=3.0, lambdaSigma1=3.0,
                 ccelec=332.0716):
        L = _lib.load()
        self._L = L
        self.at, self.ct = at, ct
        n = len(systems)
        sizes = (ctypes.c_int * (3 * n))()
        boxes = (ctypes.c_double * (3 * n))()
        counts = (ctypes.c_int * n)()
        self.counts = []
        for k, (M, box, cnt) in enumerate(systems):
            d = _dims(M)
            b = _box(box)
            for c in range(3):
                sizes[3 * k + c] = d[c]
                boxes[3 * k + c] = b[c]
            counts[k] = int(cnt)
            self.counts.append(int(cnt))
        self.nsys = n
        self.ntot = sum(self.counts)
        h = ctypes.c_void_p()
        check(L.gse_batch_create(ctypes.byref(h), _code(at), _code(ct), n, sizes, boxes, counts,
                                 float(sigma), float(kappa), float(lambdaSigma),
                                 float(lambdaSigma1), float(ccelec)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_batch_destroy(self._h)
            self._h = None

    def _s(self):
        check(self._L.gse_batch_set_stream(self._h, _stream()))

    def step(self, xyzq, force=True, potential=False, graph=False, out=None):
        """One batched step; returns dict with force (3, N), pot (N,) and
        energy (nsys,) float64, all on the device."""
        _check_xyzq(xyzq, self.ct)
        if xyzq.shape[0] != self.ntot:
            raise ValueError(f"xyzq must hold {self.ntot} particles")
        o = {} if out is None else out
        dev = xyzq.device
        if force and "force" not in o:
            o["force"] = torch.empty((3, self.ntot), dtype=self.at, device=dev)
        if potential and "pot" not in o:
            o["pot"] = torch.empty(self.ntot, dtype=self.at, device=dev)
        if "energy" not in o:
            o["energy"] = torch.empty(self.nsys, dtype=torch.float64, device=dev)
        f = o.get("force") if force else None
        args = (_ptr(xyzq), _ptr(f[0]) if f is not None else None,
                _ptr(f[1]) if f is not None else None, _ptr(f[2]) if f is not None else None,
                _ptr(o.get("pot")) if potential else None, _ptr(o["energy"]))
        self._s()
        fn = self._L.gse_batch_step_graph if graph else self._L.gse_batch_step
        check(fn(self._h, *args))
        return o

    def set_config(self, bin_edge=8, spread_block=0, gather_block=0):
        check(self._L.gse_batch_set_config(self._h, int(bin_edge), int(spread_block),
                                           int(gather_block)))

    def get_config(self):
        c = (ctypes.c_int * 3)()
        check(self._L.gse_batch_get_config(self._h, c))
        return dict(bin_edge=c[0], spread_block=c[1], gather_block=c[2])

    def autotune(self, xyzq, reps=3):
        """Times each candidate launch shape on xyzq and keeps the fastest;
        returns (config, ms per step)."""
        _check_xyzq(xyzq, self.ct)
        ms = ctypes.c_double()
        self._s()
        check(self._L.gse_batch_autotune(self._h, _ptr(xyzq), int(reps), ctypes.byref(ms)))
        return self.get_config(), ms.value


def _recip9(recip):
    r = (ctypes.c_double * 9)()
    vals = list(recip)
    if len(vals) == 3:
        vals = [1.0 / vals[0], 0, 0, 0, 1.0 / vals[1], 0, 0, 0, 1.0 / vals[2]]
    for k in range(9):
        r[k] = float(vals[k])
    return r


class CudaPMERecip:
    """CudaPMERecip<AT,CT,CT2> (cuda/CudaPMERecip.h:36-182) on the GPU: the
    SPME reciprocal comparator, order 4, 6 or 8, BOX FFT on one GPU.  recip is
    the 9-entry reciprocal box (orthorhombic) or the three box lengths."""

    def __init__(self, at, ct, nfftx, nffty, nfftz, order):
        L = _lib.load()
        self.at = _TD[_code(at)]
        self.ct = _TD[_code(ct)]
        self.dims = (nfftx, nffty, nfftz)
        self.order = order
        h = ctypes.c_void_p()
        check(L.gse_pme_recip_create(ctypes.byref(h), _code(at), _code(ct), nfftx, nffty, nfftz,
                                     order))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_pme_recip_destroy(self._h)
            self._h = None

    def _s(self):
        check(self._L.gse_pme_recip_set_stream(self._h, _stream()))

    def spread_charge(self, xyzq, ncoord, recip):
        _check_xyzq(xyzq, self.ct)
        _check_count(ncoord, xyzq)
        self._s()
        check(self._L.gse_pme_recip_spread_charge(self._h, _ptr(xyzq), ncoord, _recip9(recip)))

    def r2c_fft(self):
        self._s()
        check(self._L.gse_pme_recip_r2c_fft(self._h))

    def c2r_fft(self):
        self._s()
        check(self._L.gse_pme_recip_c2r_fft(self._h))

    def scalar_sum(self, recip, kappa, calc_energy=True, calc_virial=True):
        """Returns dict with 'energy' (1,) and 'virial' (9,) device float64."""
        o = {"energy": torch.zeros(1, dtype=torch.float64, device="cuda"),
             "virial": torch.zeros(9, dtype=torch.float64, device="cuda")}
        self._s()
        check(self._L.gse_pme_recip_scalar_sum(self._h, _recip9(recip), float(kappa),
                                               int(bool(calc_energy)), int(bool(calc_virial)),
                                               _ptr(o["energy"]), _ptr(o["virial"])))
        return o

    def gather_force(self, xyzq, ncoord, recip, stride=None):
        """dE/dr (the reference's sign) as a (3, stride) AT tensor."""
        _check_xyzq(xyzq, self.ct)
        _check_count(ncoord, xyzq)
        stride = ncoord if stride is None else stride
        f = torch.zeros((3, max(stride, 1)), dtype=self.at, device=xyzq.device)
        self._s()
        check(self._L.gse_pme_recip_gather_force(self._h, _ptr(xyzq), ncoord, _recip9(recip),
                                                 stride, _ptr(f)))
        return f

    def calc_self_energy(self, xyzq, ncoord, kappa):
        _check_xyzq(xyzq, self.ct)
        _check_count(ncoord, xyzq)
        e = torch.zeros(1, dtype=torch.float64, device=xyzq.device)
        self._s()
        check(self._L.gse_pme_recip_calc_self_energy(self._h, _ptr(xyzq), ncoord, float(kappa),
                                                     _ptr(e)))
        return e

    def grid(self):
        """Copy of the handle's grid, (Z, Y, X) of CT."""
        X, Y, Z = self.dims
        out = torch.empty((Z, Y, X), dtype=self.ct, device="cuda")
        p = ctypes.c_void_p()
        check(self._L.gse_pme_recip_get_grid(self._h, ctypes.byref(p)))
        torch.cuda.synchronize()
        check(self._L.gse_memcpy_device_to_device(_ptr(out), p, out.numel() * out.element_size()))
        return out

    def step(self, xyzq, box, kappa, force=True, virial=True, out=None):
        """spread -> r2c -> scalar_sum -> c2r -> gather: energy (1,), virial
        (9,) and forces -dE/dr (3, N)."""
        _check_xyzq(xyzq, self.ct)
        n = xyzq.shape[0]
        o = {} if out is None else out
        o.setdefault("energy", torch.empty(1, dtype=torch.float64, device=xyzq.device))
        if virial:
            o.setdefault("virial", torch.empty(9, dtype=torch.float64, device=xyzq.device))
        if force:
            o.setdefault("force", torch.empty((3, n), dtype=self.at, device=xyzq.device))
        f = o.get("force") if force else None
        self._s()
        check(self._L.gse_pme_recip_step(
            self._h, _ptr(xyzq), n, *_box(box), float(kappa), _ptr(o["energy"]),
            _ptr(o["virial"]) if virial else None, _ptr(f[0]) if f is not None else None,
            _ptr(f[1]) if f is not None else None, _ptr(f[2]) if f is not None else None))
        return o


class CpuEwaldDirect:
    """CpuEwaldDirect (CpuEwaldDirect.h:10-44) on the GPU, FP64: direct-space
    erfc pair sum within rcut (minimum image) plus the self energy."""

    def __init__(self, sigma, rcut, boxx, boxy, boxz, ccelec=332.0716):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.gse_ewald_direct_create(ctypes.byref(h), float(sigma), float(rcut), float(boxx),
                                        float(boxy), float(boxz), float(ccelec)))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_ewald_direct_destroy(self._h)
            self._h = None

    def setupExclusions(self, n_iblo14, iblo14, inb14):
        import numpy as np
        ib = np.ascontiguousarray(iblo14, dtype=np.int32)
        nb = np.ascontiguousarray(inb14 if len(inb14) else [0], dtype=np.int32)
        if ib.size < n_iblo14:
            raise ValueError("iblo14 needs n_iblo14 entries")
        self._excl = (ib, nb)
        ip = ctypes.POINTER(ctypes.c_int)
        check(self._L.gse_ewald_direct_setup_exclusions(self._h, int(n_iblo14),
                                                        ib.ctypes.data_as(ip),
                                                        nb.ctypes.data_as(ip)))

    def calcForceEnergy(self, n, xyzq, forceX, forceY, forceZ):
        """Writes dE/dr (the reference's sign) into the three float64 device
        arrays and returns E_direct + E_self (synchronises)."""
        _check_xyzq(xyzq, torch.float64)
        _check_count(n, xyzq)
        for f in (forceX, forceY, forceZ):
            if f.dtype != torch.float64 or f.numel() < n:
                raise ValueError("force arrays must be float64 with n entries")
        e = torch.empty(1, dtype=torch.float64, device=xyzq.device)
        check(self._L.gse_ewald_direct_set_stream(self._h, _stream()))
        check(self._L.gse_ewald_direct_calc_force_energy(self._h, n, _ptr(xyzq), _ptr(forceX),
                                                         _ptr(forceY), _ptr(forceZ), _ptr(e)))
        return float(e.cpu()[0])


class CpuEwaldRecip:
    """CpuEwaldRecip (CpuEwaldRecip.h:12-46) on the GPU, FP64: the exact Ewald
    k-space sum over 0 < |m| <= kcut."""

    def __init__(self, sigma, kcut, boxx, boxy, boxz, ccelec=332.0716):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.gse_ewald_recip_create(ctypes.byref(h), float(sigma), int(kcut), float(boxx),
                                       float(boxy), float(boxz), float(ccelec)))
        self._h = h
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.gse_ewald_recip_destroy(self._h)
            self._h = None

    def calcForceEnergy(self, n, xyzq, forceX=None, forceY=None, forceZ=None):
        """Writes the forces -dE/dr into the float64 device arrays (all None:
        energy only) and returns the energy (synchronises)."""
        _check_xyzq(xyzq, torch.float64)
        _check_count(n, xyzq)
        fs = (forceX, forceY, forceZ)
        if any(f is not None for f in fs):
            for f in fs:
                if f is None or f.dtype != torch.float64 or f.numel() < n:
                    raise ValueError("force arrays must be float64 with n entries")
        e = torch.empty(1, dtype=torch.float64, device=xyzq.device)
        check(self._L.gse_ewald_recip_set_stream(self._h, _stream()))
        check(self._L.gse_ewald_recip_calc_force_energy(self._h, n, _ptr(xyzq), *(
            _ptr(f) if f is not None else None for f in fs), _ptr(e)))
        return float(e.cpu()[0])


class _GSEMethod:
    """Shared parts of the GSE-r and GSE-u mirrors (prefix: C-ABI name stem)."""

    _prefix = ""

    def _fn(self, name):
        return getattr(self._L, f"gse_{self._prefix}_{name}")

    def __del__(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = None

    def _s(self):
        check(self._fn("set_stream")(self._h, _stream()))

    def spreadCharge1(self, numCoord, xyzq):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        self._s()
        check(self._fn("spread_charge1")(self._h, numCoord, _ptr(xyzq)))

    def solvePoisson(self):
        self._s()
        check(self._fn("solve_poisson")(self._h))

    def calculateEnergy(self):
        e = ctypes.c_double()
        self._s()
        check(self._fn("calculate_energy")(self._h, ctypes.byref(e)))
        return e.value

    def interpolateForce(self, numCoord, xyzq, forceX, forceY, forceZ):
        _check_xyzq(xyzq, self.ct)
        _check_count(numCoord, xyzq)
        for f in (forceX, forceY, forceZ):
            if f.dtype != self.at or f.numel() < numCoord:
                raise ValueError("force arrays must be AT with numCoord entries")
        self._s()
        check(self._fn("interpolate_force")(self._h, numCoord, _ptr(xyzq), _ptr(forceX),
                                            _ptr(forceY), _ptr(forceZ)))

    def _copy_grid(self, which):
        X, Y, Z = self.dims
        out = torch.empty((Z, Y, X), dtype=self.ct, device="cuda")
        self._s()
        check(self._fn("copy_grid")(self._h, which, _ptr(out)))
        return out

    def step(self, xyzq, force=True, potential=False, cube=False, out=None):
        """Fused step: energy (device float64) and the requested outputs."""
        _check_xyzq(xyzq, self.ct)
        n = xyzq.shape[0]
        dev = xyzq.device
        o = {} if out is None else out
        d = _lib.StepDesc()
        if force:
            o.setdefault("force", torch.empty((3, n), dtype=self.at, device=dev))
            d.forceX, d.forceY, d.forceZ = (_ptr(o["force"][k]) for k in range(3))
        if potential:
            o.setdefault("pot", torch.empty(n, dtype=self.at, device=dev))
            d.potential = _ptr(o["pot"])
        if cube:
            o.setdefault("force_cube", torch.empty((3, n), dtype=self.at, device=dev))
            d.cubeX, d.cubeY, d.cubeZ = (_ptr(o["force_cube"][k]) for k in range(3))
        o.setdefault("energy", torch.empty(1, dtype=torch.float64, device=dev))
        d.energy = _ptr(o["energy"])
        self._s()
        check(self._fn("step")(self._h, n, _ptr(xyzq), ctypes.byref(d)))
        return o


class CpuGSEr(_GSEMethod):
    """CpuGSEr<AT,CT> (CpuGSEr.h:18-37) on the GPU: sigma1 spread, on-grid
    sigma2 convolution, the HERMITIAN-stencil Poisson equation solved exactly
    (zero-mean potential), energy from the sigma1 grid."""

    _prefix = "gser"

    def __init__(self, at, ct, sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma, lambdaSigma1,
                 boxx, boxy, boxz, ccelec=332.0716):
        L = _lib.load()
        self._L = L
        self.at = _TD[_code(at)]
        self.ct = _TD[_code(ct)]
        self.dims = (sizeX, sizeY, sizeZ)
        h = ctypes.c_void_p()
        check(L.gse_gser_create(ctypes.byref(h), _code(at), _code(ct), sizeX, sizeY, sizeZ,
                                sigma, kappa, lambdaSigma, lambdaSigma1, boxx, boxy, boxz,
                                ccelec))
        self._h = h

    def spreadCharge2(self):
        self._s()
        check(self._L.gse_gser_spread_charge2(self._h))

    def getRho(self):
        return self._copy_grid(0)

    def getRhoS(self):
        return self._copy_grid(1)

    def getPhi(self):
        return self._copy_grid(2)


class CpuGSEu(_GSEMethod):
    """CpuGSEu<AT,CT> (CpuGSEu.h:18-42) on the GPU: the u-series of N on-grid
    convolutions with widths sigma_j = sqrt(sigma^2 b^(2j) - 2 sigma1^2)."""

    _prefix = "gseu"

    def __init__(self, at, ct, N, b, sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma,
                 lambdaSigma1, boxx, boxy, boxz, ccelec=332.0716):
        L = _lib.load()
        self._L = L
        self.at = _TD[_code(at)]
        self.ct = _TD[_code(ct)]
        self.dims = (sizeX, sizeY, sizeZ)
        h = ctypes.c_void_p()
        check(L.gse_gseu_create(ctypes.byref(h), _code(at), _code(ct), int(N), float(b), sizeX,
                                sizeY, sizeZ, sigma, kappa, lambdaSigma, lambdaSigma1, boxx, boxy,
                                boxz, ccelec))
        self._h = h

    def getRhoS(self):
        return self._copy_grid(0)

    def getPhi(self):
        return self._copy_grid(1)
