"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/gse_hip.h declares, and validates arguments before any device work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gse_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"GSE_API\s+[^;(]*?\b(gse_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from gse_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgse_hip.so not built")
    return _lib.load()


def test_header_and_binding_agree():
    from gse_amd import _lib
    assert declared() == sorted(_lib.EXPORTED)


def test_every_declared_symbol_is_exported(lib):
    for name in declared():
        assert hasattr(lib, name), name


def test_status_strings(lib):
    assert lib.gse_status_string(0) == b"success"
    assert lib.gse_status_string(1) == b"invalid value"
    assert lib.gse_version() >= 10000


def test_argument_validation_without_device(lib):
    h = ctypes.c_void_p()
    assert lib.gse_gauss_charge_create(ctypes.byref(h), 0, 1, 8, 8, 8) == 1  # (f32, f64)
    assert b"pair" in lib.gse_last_error()
    assert lib.gse_gauss_charge_create(ctypes.byref(h), 1, 1, 0, 8, 8) == 1
    assert lib.gse_convolution_create(ctypes.byref(h), 5, 8, 8, 8) == 1
    assert lib.gse_gsek_create(ctypes.byref(h), 1, 1, 8, 8, 8, 2.0, 0.6, 3.0, 3.0,
                               8.0, 8.0, 8.0, 332.0716) == 1  # kappa > 0.5
    assert lib.gse_set_tiling(-1, 0) == 1
    assert lib.gse_set_tiling(0, 0) == 0
    # handle-free queries of a valid CpuGaussCharge touch no device memory
    assert lib.gse_gauss_charge_create(ctypes.byref(h), 1, 1, 20, 20, 20) == 0
    n = ctypes.c_int()
    assert lib.gse_gauss_charge_calc_support_size(h, 5.6921, 20.0, 20.0, 20.0, ctypes.byref(n)) == 0
    from oracle import oracle
    assert n.value == oracle.support_size((20, 20, 20), 5.6921, (20.0,) * 3)
    assert lib.gse_gauss_charge_destroy(h) == 0


def test_tuning_knobs_round_trip_and_reject_bad_values(lib):
    """Every named launch-shape knob reads back what was set, rejects values
    outside its documented range with a message naming it, and leaves its
    value unchanged on rejection.  Setting knobs touches no device."""
    from gse_amd import api
    keys = api.get_tuning()
    assert set(keys) == set(api.TUNING_KEYS)
    good = {"bin_edge": 6, "brick_budget": 32768, "spread_block": 256, "gather_block": 256,
            "spread_stage": 8192, "gather_stage": 16384, "spread_lds32": 0, "spread_cols": 1,
            "spread_flush": 0, "reduce_tile": 0, "reduce_nt": 1, "reduce_tile2": 0, "bin_agg": 2,
            "fused_kz": 0, "spread_mma": 0, "zpass_cols": 16, "gather_group": 16,
            "zpass4": 0, "xy_plane": 0, "zpass4_cols": 32,
            "bin_coarse_block": 512, "ypass4": 0, "bin_order": 0}
    bad = {"bin_edge": 0, "spread_block": 384, "gather_block": 1024, "spread_stage": 512,
           "gather_stage": 200000, "spread_lds32": 2, "reduce_tile": -1, "reduce_nt": 2,
           "reduce_tile2": 5, "bin_agg": 3, "fused_kz": 2, "spread_mma": -1, "zpass_cols": 5,
           "gather_group": 8, "zpass4": 2,
           "xy_plane": -1, "zpass4_cols": 8,
           "bin_coarse_block": 128, "ypass4": 2, "bin_order": 3}
    assert set(good) == set(keys)
    try:
        for k, v in good.items():
            api.set_tuning(**{k: v})
            assert api.get_tuning()[k] == v, k
        for k, v in bad.items():
            before = api.get_tuning()[k]
            assert lib.gse_set_tuning_param(k.encode(), v) == 1, k
            assert k.encode() in lib.gse_last_error(), k
            assert api.get_tuning()[k] == before, k
        assert lib.gse_set_tuning_param(b"no_such_knob", 1) == 1
    finally:
        api.set_tuning(**keys)
    assert api.get_tuning() == keys
