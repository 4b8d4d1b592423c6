"""ctypes loader of gse_amd/libgse_hip.so (the C-ABI in include/gse_hip.h).

torch is imported before the library: torch ships its own HIP runtime with
the same soname, and loading it first makes this library bind to that one
runtime, so device pointers and streams from torch are valid here.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgse_hip.so")

GSE_FLOAT32 = 0
GSE_FLOAT64 = 1
GREENS_GSE = 0
GREENS_G2 = 1
GREENS_HERMITIAN = 2

GATHER_SPHERE_FORCE = 1
GATHER_CUBE_FORCE = 2
GATHER_POTENTIAL = 4
GATHER_VIRIAL = 8

LAYOUT_FIELDS = 12
# conversion kernel of the last spread (gse_*_get_layout field 8)
REDUCE_PATHS = {-1: "none", 0: "atomic", 1: "generic", 2: "tile4", 3: "tile4_nt", 4: "tile2",
                5: "tile"}

# k-space form of an FFT solve (gse_*_get_layout field 10, gse_fft_solver_get_form)
KSPACE_FORMS = {-1: "none", 0: "3d", 1: "fused", 2: "fused4", 3: "fused/xy", 4: "fused4/xy",
                5: "fused/y4", 6: "fused4/y4"}

_lib = None

V = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
PD = ctypes.POINTER(ctypes.c_double)
PI = ctypes.POINTER(ctypes.c_int)
PV = ctypes.POINTER(ctypes.c_void_p)


class GatherDesc(ctypes.Structure):
    _fields_ = [("what", I), ("scale", D), ("forceX", V), ("forceY", V), ("forceZ", V),
                ("cubeX", V), ("cubeY", V), ("cubeZ", V), ("potential", V), ("virial", V)]


class StepDesc(ctypes.Structure):
    _fields_ = [("mode", I), ("forceX", V), ("forceY", V), ("forceZ", V), ("potential", V),
                ("cubeX", V), ("cubeY", V), ("cubeZ", V), ("energy", V), ("virial", V)]


# name: argtypes (restype is always the int status unless listed in _RET)
_SIGS = {
    "gse_status_string": [I],
    "gse_last_error": [],
    "gse_version": [],
    "gse_set_tiling": [I, I],
    "gse_set_tuning_param": [ctypes.c_char_p, I],
    "gse_get_tuning_param": [ctypes.c_char_p, PI],
    "gse_device_malloc": [PV, ctypes.c_size_t],
    "gse_device_free": [V],
    "gse_memcpy_to_device": [V, V, ctypes.c_size_t],
    "gse_memcpy_to_host": [V, V, ctypes.c_size_t],
    "gse_memcpy_device_to_device": [V, V, ctypes.c_size_t],
    "gse_device_clear": [V, ctypes.c_size_t],
    "gse_device_synchronize": [],
    "gse_gauss_charge_create": [PV, I, I, I, I, I],
    "gse_gauss_charge_destroy": [V],
    "gse_gauss_charge_set_stream": [V, V],
    "gse_gauss_charge_calc_support_size": [V, D, D, D, D, PI],
    "gse_gauss_charge_spread_charge_to_grid": [V, D, D, I, V, D, D, D, V, I],
    "gse_gauss_charge_spread_charge_to_grid_noalloc": [V, D, D, I, V, D, D, D, V, I],
    "gse_gauss_charge_spread_charge_on_grid": [V, D, D, D, D, D, V, V],
    "gse_gauss_charge_interpolate_force": [V, D, D, I, V, D, D, D, V, V, V, V],
    "gse_gauss_charge_calc_electric_field_on_grid": [V, D, D, D, V, V, V, V],
    "gse_gauss_charge_calc_dipole_sum": [V, D, D, D, V, PD],
    "gse_gauss_charge_gather": [V, D, D, I, V, D, D, D, V, ctypes.POINTER(GatherDesc)],
    "gse_gauss_charge_invalid_count": [V, PI],
    "gse_gauss_charge_get_layout": [V, PI],
    "gse_convolution_create": [PV, I, I, I, I],
    "gse_convolution_destroy": [V],
    "gse_convolution_set_stream": [V, V],
    "gse_convolution_setup_kernel": [V, I, I, I, V, V, V],
    "gse_convolution_run": [V, V, V],
    "gse_convolution_run_xyz": [V, V, V],
    "gse_fft_solver_create": [PV, I, I, I, I, I],
    "gse_fft_solver_destroy": [V],
    "gse_fft_solver_set_stream": [V, V],
    "gse_fft_solver_run": [V, V, V, D, D, D, D],
    "gse_fft_solver_r2c": [V, V, V],
    "gse_fft_solver_get_form": [V, PI],
    "gse_fft_solver_c2r": [V, V, V],
    "gse_greens_func_mul": [I, I, I, I, I, D, D, D, D, V, V],
    "gse_gsek_create": [PV, I, I, I, I, I, D, D, D, D, D, D, D, D],
    "gse_gsek_destroy": [V],
    "gse_gsek_set_stream": [V, V],
    "gse_gsek_set_box_size": [V, D, D, D],
    "gse_gsek_get_params": [V, PD, PD, PD],
    "gse_gsek_spread_charge1": [V, I, V],
    "gse_gsek_spread_charge2": [V, V],
    "gse_gsek_solve_poisson": [V],
    "gse_gsek_calculate_energy": [V, PD],
    "gse_gsek_interpolate_force": [V, I, V, V, V, V],
    "gse_gsek_get_layout": [V, PI],
    "gse_gsek_get_rho_s": [V, PV],
    "gse_gsek_get_phi": [V, PV],
    "gse_gsek_copy_grid": [V, I, V],
    "gse_gsek_step": [V, I, V, ctypes.POINTER(StepDesc)],
    "gse_gsek_adjoint": [V, I, V, V, V, V, V, V, V],
    "gse_gsek_set_profiling": [V, I],
    "gse_gsek_stage_times": [V, PD, PI],
    "gse_gsek_corrections": [V, I, V, V],
    "gse_batch_create": [PV, I, I, I, PI, PD, PI, D, D, D, D, D],
    "gse_batch_destroy": [V],
    "gse_batch_set_stream": [V, V],
    "gse_batch_step": [V, V, V, V, V, V, V],
    "gse_batch_step_graph": [V, V, V, V, V, V, V],
    "gse_batch_set_config": [V, I, I, I],
    "gse_batch_get_config": [V, PI],
    "gse_batch_autotune": [V, V, I, PD],
    "gse_pme_recip_create": [PV, I, I, I, I, I, I],
    "gse_pme_recip_destroy": [V],
    "gse_pme_recip_set_stream": [V, V],
    "gse_pme_recip_spread_charge": [V, V, I, PD],
    "gse_pme_recip_r2c_fft": [V],
    "gse_pme_recip_c2r_fft": [V],
    "gse_pme_recip_scalar_sum": [V, PD, D, I, I, V, V],
    "gse_pme_recip_gather_force": [V, V, I, PD, I, V],
    "gse_pme_recip_calc_self_energy": [V, V, I, D, V],
    "gse_pme_recip_step": [V, V, I, D, D, D, D, V, V, V, V, V],
    "gse_pme_recip_get_grid": [V, PV],
    "gse_ewald_direct_create": [PV, D, D, D, D, D, D],
    "gse_ewald_direct_destroy": [V],
    "gse_ewald_direct_set_stream": [V, V],
    "gse_ewald_direct_setup_exclusions": [V, I, PI, PI],
    "gse_ewald_direct_calc_force_energy": [V, I, V, V, V, V, V],
    "gse_gser_create": [PV, I, I, I, I, I, D, D, D, D, D, D, D, D],
    "gse_gser_destroy": [V],
    "gse_gser_set_stream": [V, V],
    "gse_gser_spread_charge1": [V, I, V],
    "gse_gser_spread_charge2": [V],
    "gse_gser_solve_poisson": [V],
    "gse_gser_calculate_energy": [V, PD],
    "gse_gser_interpolate_force": [V, I, V, V, V, V],
    "gse_gser_copy_grid": [V, I, V],
    "gse_gser_step": [V, I, V, ctypes.POINTER(StepDesc)],
    "gse_gseu_create": [PV, I, I, I, D, I, I, I, D, D, D, D, D, D, D, D],
    "gse_gseu_destroy": [V],
    "gse_gseu_set_stream": [V, V],
    "gse_gseu_spread_charge1": [V, I, V],
    "gse_gseu_solve_poisson": [V],
    "gse_gseu_calculate_energy": [V, PD],
    "gse_gseu_interpolate_force": [V, I, V, V, V, V],
    "gse_gseu_copy_grid": [V, I, V],
    "gse_gseu_step": [V, I, V, ctypes.POINTER(StepDesc)],
    "gse_ewald_recip_create": [PV, D, I, D, D, D, D],
    "gse_ewald_recip_destroy": [V],
    "gse_ewald_recip_set_stream": [V, V],
    "gse_ewald_recip_calc_force_energy": [V, I, V, V, V, V, V],
}
_RET = {"gse_status_string": ctypes.c_char_p, "gse_last_error": ctypes.c_char_p,
        "gse_version": I}

EXPORTED = tuple(_SIGS)


class GSEError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"gse status {status}: {message}")
        self.status = status


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (shares its HIP runtime, see module docstring)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " or `make -C gse_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RET.get(name, I)
    _lib = L
    return L


def check(status):
    if status != 0:
        msg = _lib.gse_last_error().decode() if _lib is not None else ""
        raise GSEError(status, msg)
