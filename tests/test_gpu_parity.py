"""HIP field kernel against the golden vectors and the C oracle.

Bar: escape counts, NaN mask, test1/test2 planes and all RenderStats
counters bit-exact; float32 field within one ulp (x87 long double in the
reference's final logarithm); phase/DE within 1e-9 relative."""
import ctypes
import sys

import numpy as np
import pytest

import golden_io
import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.case_names())
def test_hip_matches_golden(gpu_product, name):
    print(f"[gpu] golden case {name}", flush=True)
    golden = golden_io.load(name)
    field, planes, stats = golden_io.render_abi(gpu_product, golden["meta"])
    golden_io.assert_hip_parity(golden, field, planes, stats)


@pytest.mark.parametrize("name", ["c1_series", "c2_kalles", "bla_1e25"])
def test_hip_field_without_planes_equals_with_planes(gpu_product, name):
    print(f"[gpu] planes/no-planes {name}", flush=True)
    golden = golden_io.load(name)
    meta = golden["meta"]
    opts = native.RenderOptions.defaults(backend=2, **meta["options"])
    ref = native.AbiReference(gpu_product, meta["x"], meta["y"], meta["ref_zoom"], meta["max_iter"],
                              meta["bits"], meta["image_series_order"], opts)
    a, _, _ = ref.render(meta["width"], meta["height"], meta["zoom"], meta["max_iter"],
                         meta["series_order"], meta["series_block"], opts)
    b, _, _ = ref.render(meta["width"], meta["height"], meta["zoom"], meta["max_iter"],
                         meta["series_order"], meta["series_block"], opts, planes=True)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_hip_full_640x360_matches_oracle(gpu_product, oracle):
    """The 640x360 zoom-1e12 workload at full size against the oracle."""
    print("[gpu] 640x360 1e12 full frame", flush=True)
    meta = dict(golden_io.load("c1_series")["meta"], width=640, height=360)
    f1, p1, s1 = golden_io.render_abi(gpu_product, meta)
    f2, p2, s2 = golden_io.render_oracle(oracle, meta)
    golden = {"field": f2, "orbit_iteration": p2.orbit_iteration, "stats": s2,
              **{k: getattr(p2, k) for k in ("phase", "de_x", "de_y", "test1", "test2")}}
    golden_io.assert_hip_parity(golden, f1, p1, s1)


def test_hip_ragged_and_tiny_fields(gpu_product, oracle):
    """Sizes that do not fill a 16x16 tile, and a single pixel."""
    base = golden_io.load("c1_series")["meta"]
    for w, h in [(1, 1), (17, 3), (33, 31)]:
        print(f"[gpu] ragged {w}x{h}", flush=True)
        meta = dict(base, width=w, height=h)
        f1, p1, s1 = golden_io.render_abi(gpu_product, meta)
        f2, p2, s2 = golden_io.render_oracle(oracle, meta)
        golden = {"field": f2, "orbit_iteration": p2.orbit_iteration, "stats": s2,
                  **{k: getattr(p2, k) for k in ("phase", "de_x", "de_y", "test1", "test2")}}
        golden_io.assert_hip_parity(golden, f1, p1, s1)


def test_time_budget_marks_unfinished_pixels(gpu_product):
    """A 1 ms budget on a ~100k-iteration field leaves NaN pixels
    (renderer.cpp:11872-11890; TNA:638-675)."""
    print("[gpu] time budget", flush=True)
    meta = golden_io.load("recovery_1e100")["meta"]
    opts = native.RenderOptions.defaults(backend=2, time_budget_ms=1, **meta["options"])
    ref = native.AbiReference(gpu_product, meta["x"], meta["y"], meta["ref_zoom"], meta["max_iter"],
                              meta["bits"], meta["image_series_order"], opts)
    gpu_product.fractal_set_stats_enabled(1)
    field, _, stats = ref.render(64, 36, meta["zoom"], meta["max_iter"], meta["series_order"],
                                 meta["series_block"], opts, stats=True)
    assert np.isnan(field).any()
    assert stats[20] > 0 and stats[19] >= stats[20] // 2


def test_device_pointer_entry_matches_host_entry(gpu_product):
    """fractal_hip_render_reference_device writes the same field into HBM."""
    print("[gpu] device-pointer entry", flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    golden = golden_io.load("c2_kalles")
    meta = golden["meta"]
    opts = native.RenderOptions.defaults(backend=2, **meta["options"])
    ref = native.AbiReference(gpu_product, meta["x"], meta["y"], meta["ref_zoom"], meta["max_iter"],
                              meta["bits"], meta["image_series_order"], opts)
    n = meta["width"] * meta["height"]
    dptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dptr), ctypes.c_size_t(n * 4)) == 0
    try:
        gpu_product.fractal_hip_render_reference_device.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.POINTER(native.RenderOptions), ctypes.c_void_p]
        status = gpu_product.fractal_hip_render_reference_device(
            dptr, meta["width"], meta["height"], meta["zoom"].encode(), ctypes.c_void_p(ref.handle),
            meta["max_iter"], meta["series_order"], meta["series_block"], ctypes.byref(opts), None)
        assert status == 0, gpu_product.fractal_last_error()
        host = np.empty(n, dtype=np.float32)
        assert hip.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), dptr, ctypes.c_size_t(n * 4), 2) == 0
    finally:
        hip.hipFree(dptr)
    np.testing.assert_array_equal(host.view(np.uint32), golden["field"].reshape(-1).view(np.uint32))
