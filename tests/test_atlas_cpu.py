"""Raw KFP atlas colour pass on the CPU: the C oracle against the
reference's golden vectors (bit-exact RGB), an invariant the reference
holds (centred tiles decode to the same frame), and the library's argument
checks, which run before any device work."""
import ctypes

import numpy as np
import pytest

import atlas_io
import native


@pytest.fixture(scope="module")
def oracle_atlas(oracle):
    oracle.fo_atlas_colourise_kfp_raw.argtypes = (
        atlas_io.ATLAS_ARGS[:15] + [ctypes.POINTER(native.KfpOptions), ctypes.POINTER(ctypes.c_uint8),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int])
    oracle.fo_atlas_colourise_kfp_raw.restype = ctypes.c_int
    return oracle


def oracle_call(oracle, case, parent, child, lut, opts):
    ow, oh = case["output"]
    pw, ph, pmax = case["parent"]
    cw, ch, cmax = case["child"] if child is not None else (0, 0, 0)
    out = np.zeros((oh, ow, 3), dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    p = np.ascontiguousarray(parent, dtype=np.float32)
    c = np.ascontiguousarray(child, dtype=np.float32) if child is not None else None
    status = oracle.fo_atlas_colourise_kfp_raw(
        p.ctypes.data_as(atlas_io.c_float_p), pw, ph, pmax,
        c.ctypes.data_as(atlas_io.c_float_p) if c is not None else None, cw, ch, cmax,
        out.ctypes.data_as(atlas_io.c_uint8_p), ow, oh, case["parent_zoom"], case["child_fraction"],
        case["child_zoom"], case["max_iter"], ctypes.byref(opts), lut.ctypes.data_as(atlas_io.c_uint8_p),
        lut.shape[0], 0, err, 256)
    assert status == 0, err.value
    return out


@pytest.mark.parametrize("name", atlas_io.case_names())
def test_oracle_atlas_matches_reference_golden(oracle_atlas, name):
    g = atlas_io.load(name)
    parent, child, lut, opts = atlas_io.case_inputs(oracle_atlas, g["meta"])
    rgb = oracle_call(oracle_atlas, g["meta"], parent, child, lut, opts)
    np.testing.assert_array_equal(rgb, g["rgb"])


def test_centered_tiles_decode_to_nearly_the_same_frame():
    """Tiles stored relative to their own caps (field_bias > 0) decode to the
    frame the uncentred tiles give (renderer.cpp:18026-18035), up to the
    float rounding of the stored offsets: at most one level on under 2% of
    the pixels."""
    a = atlas_io.load("atlas_centered")["rgb"].astype(int)
    b = atlas_io.load("atlas_child_half")["rgb"].astype(int)
    diff = np.abs(a - b).max(axis=2)
    assert diff.max() <= 1 and (diff > 0).mean() < 0.02


def test_product_atlas_checks_arguments_before_the_device(product, oracle):
    fn = atlas_io.bind_atlas(product)
    case = atlas_io.CASES[0]
    parent, child, lut, opts = atlas_io.case_inputs(oracle, case)
    bad = [dict(case, child_fraction=1.5), dict(case, parent_zoom=0.0), dict(case, child_zoom=float("nan")),
           dict(case, output=[0, 10])]
    for b in bad:
        status, _ = atlas_io.atlas_call(fn, b, parent, child, lut, opts)
        assert status == 1
        assert b"invalid native raw KFP atlas dimensions or controls" in product.fractal_last_error()
    status, _ = atlas_io.atlas_call(fn, dict(case, child_fraction=0.0), parent, child, lut, opts)
    assert status == 1 and b"positive visible fraction" in product.fractal_last_error()
    status, _ = atlas_io.atlas_call(fn, dict(case, child=[0, 0, 0]), parent, child, lut, opts)
    assert status == 1 and b"child tile" in product.fractal_last_error()
    biased = native.KfpOptions.make(**dict(case["options"], field_bias=7000.0))
    status, _ = atlas_io.atlas_call(fn, case, parent, child, lut, biased)
    assert status == 1 and b"field bias exceeds" in product.fractal_last_error()
