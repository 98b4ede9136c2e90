"""fractal_hip_colourise_kfp_device_noalloc: the KFP colour pass with its
scratch (LUT staging, colour-method-4 range reduction) in the library's
fixed scratch array instead of device allocations.  Two calls in a row,
the second smaller and with a different method and LUT, must equal the
allocating device entry bit for bit, and device memory must not change
across calls."""
import ctypes

import numpy as np
import pytest

import native
from kfp_cases import KALLES_DEFAULT, synthetic_field


def _colour(lib, entry, hip, field, max_iter, opts, lut):
    h, w = field.shape
    d_field = hip.malloc(h * w * 4)
    d_rgb = hip.malloc(h * w * 3)
    try:
        hip.to_device(d_field, np.ascontiguousarray(field, dtype=np.float32))
        table = np.ascontiguousarray(lut, dtype=np.uint8)
        status = getattr(lib, entry)(d_field, d_rgb, w, h, max_iter, ctypes.byref(opts),
                                     table.ctypes.data_as(native.c_uint8_p), table.shape[0])
        assert status == 0, lib.fractal_last_error()
        rgb = np.empty((h, w, 3), dtype=np.uint8)
        hip.to_host(rgb, d_rgb)
    finally:
        hip.free(d_field)
        hip.free(d_rgb)
    return rgb


def _free_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.lib.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.gpu
def test_noalloc_kfp_equals_device_entry_on_two_calls(gpu_product, oracle):
    import fav_hip
    hip = fav_hip.HipRuntime()
    cases = [
        ("1920x1080 method 4 (range reduction), 1024-entry LUT", 1920, 1080, 4, 1024),
        ("640x360 method 7, 256-entry LUT", 640, 360, 7, 256),
    ]
    for label, w, h, method, entries in cases:
        print(f"[gpu] noalloc kfp {label}", flush=True)
        field = synthetic_field(w, h, 20000, seed=11 + method)
        opts = native.KfpOptions.make(**dict(KALLES_DEFAULT, color_method=method))
        lut = native.palette_lut_oracle(oracle, [(255, 255, 255), (128, 0, 64), (0, 0, 255), (0, 255, 0)], entries)
        expected = _colour(gpu_product, "fractal_hip_colourise_kfp_device", hip, field, 20000, opts, lut)
        got = _colour(gpu_product, "fractal_hip_colourise_kfp_device_noalloc", hip, field, 20000, opts, lut)
        np.testing.assert_array_equal(got, expected, err_msg=label)


@pytest.mark.gpu
def test_noalloc_kfp_allocates_nothing_and_rejects_oversize(gpu_product, oracle):
    import fav_hip
    hip = fav_hip.HipRuntime()
    print("[gpu] noalloc kfp: device memory across calls", flush=True)
    field = synthetic_field(512, 288, 20000, seed=5)
    opts = native.KfpOptions.make(**dict(KALLES_DEFAULT, color_method=4))
    lut = native.palette_lut_oracle(oracle, [(255, 0, 0), (0, 0, 255)], 4096)
    h, w = field.shape
    d_field = hip.malloc(h * w * 4)
    d_rgb = hip.malloc(h * w * 3)
    try:
        hip.to_device(d_field, np.ascontiguousarray(field, dtype=np.float32))
        table = np.ascontiguousarray(lut, dtype=np.uint8)

        def call(entries):
            return gpu_product.fractal_hip_colourise_kfp_device_noalloc(
                d_field, d_rgb, w, h, 20000, ctypes.byref(opts), table.ctypes.data_as(native.c_uint8_p), entries)

        assert call(table.shape[0]) == 0, gpu_product.fractal_last_error()  # set-up on this device
        before = _free_bytes(hip)
        for entries in (4096, 1024, 2):
            assert call(entries) == 0, gpu_product.fractal_last_error()
        assert _free_bytes(hip) == before
        big = np.zeros((65537, 3), dtype=np.uint8)
        status = gpu_product.fractal_hip_colourise_kfp_device_noalloc(
            d_field, d_rgb, w, h, 20000, ctypes.byref(opts), big.ctypes.data_as(native.c_uint8_p), big.shape[0])
        assert status == 1
        assert b"invalid" in gpu_product.fractal_last_error()
    finally:
        hip.free(d_field)
        hip.free(d_rgb)
