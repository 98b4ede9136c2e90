"""HIP KFP colour pass against the reference's RGB and the C oracle.

Bar: every channel within one 8-bit level (the device's log/sin/pow/atan2
may round differently from the host libm right at a quantisation step);
the exact-match fraction is reported."""
import ctypes

import numpy as np
import pytest

import golden_io
import native
from kfp_cases import KALLES_DEFAULT, synthetic_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.kfp_case_names())
def test_hip_kfp_matches_reference_golden(gpu_product, name):
    print(f"[gpu] kfp golden {name}", flush=True)
    g = golden_io.load_kfp(name)
    rgb = native.colourise_abi(gpu_product, g["field"], g["max_iter"], g["options"], g["lut"])
    golden_io.assert_rgb_within_one_level(rgb, g["rgb"], name)
    exact = float(np.mean(rgb == g["rgb"]))
    print(f"[gpu] kfp {name}: exact channel fraction {exact:.6f}", flush=True)
    assert exact > 0.999


@pytest.mark.parametrize("entry", ["fractal_colourise_kfp_opencl", "fractal_colourise_kfp_precise"])
def test_hip_kfp_entry_aliases(gpu_product, entry):
    print(f"[gpu] kfp alias {entry}", flush=True)
    g = golden_io.load_kfp("kalles_default_bla")
    a = native.colourise_abi(gpu_product, g["field"], g["max_iter"], g["options"], g["lut"])
    b = native.colourise_abi(gpu_product, g["field"], g["max_iter"], g["options"], g["lut"], entry=entry)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("method", [0, 4, 7])
def test_hip_kfp_full_hd_matches_oracle(gpu_product, oracle, method):
    """A 1920x1080 synthetic field (ramps, interior blocks, NaN glitches)."""
    print(f"[gpu] kfp 1920x1080 method {method}", flush=True)
    field = synthetic_field(1920, 1080, 20000, seed=3 + method)
    opts = native.KfpOptions.make(**dict(KALLES_DEFAULT, color_method=method))
    lut = native.palette_lut_oracle(oracle, [(255, 255, 255), (128, 0, 64), (0, 0, 255), (0, 255, 0)], 1024)
    rgb = native.colourise_abi(gpu_product, field, 20000, opts, lut)
    expected = native.colourise_oracle(oracle, field, 20000, opts, lut)
    golden_io.assert_rgb_within_one_level(rgb, expected, f"method {method}")


def test_hip_kfp_device_pointers_equal_host_entry(gpu_product):
    """fractal_hip_colourise_kfp_device on HBM-resident field and RGB."""
    import fav_hip
    print("[gpu] kfp device-pointer entry", flush=True)
    g = golden_io.load_kfp("kalles_default_c1")
    h, w = g["field"].shape
    host = native.colourise_abi(gpu_product, g["field"], g["max_iter"], g["options"], g["lut"])
    hip = fav_hip.HipRuntime()
    d_field = hip.malloc(h * w * 4)
    d_rgb = hip.malloc(h * w * 3)
    try:
        hip.to_device(d_field, np.ascontiguousarray(g["field"], dtype=np.float32))
        lut = np.ascontiguousarray(g["lut"], dtype=np.uint8)
        status = gpu_product.fractal_hip_colourise_kfp_device(
            d_field, d_rgb, w, h, g["max_iter"], ctypes.byref(g["options"]),
            lut.ctypes.data_as(native.c_uint8_p), lut.shape[0])
        assert status == 0, gpu_product.fractal_last_error()
        rgb = np.empty((h, w, 3), dtype=np.uint8)
        hip.to_host(rgb, d_rgb)
    finally:
        hip.free(d_field)
        hip.free(d_rgb)
    np.testing.assert_array_equal(rgb, host)
