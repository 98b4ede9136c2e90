"""The plane-aware KFP colour pass (kfp_core.h compiled for the CPU by the
test harness) against the fixtures the reference's renderer.cpp wrote:
identical RGB, since both use the host libm; plus the HIP library's
argument checks for planes."""
import ctypes

import numpy as np
import pytest

import host_check
import kfp_planes_io
import native
from golden_io import GOLDEN
from kfp_planes_cases import CASES


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as g:
        return g["rgb"]


@pytest.fixture(scope="module")
def harness():
    if not host_check.HOST_CHECK_SO.exists():
        pytest.fail("libfav_host_check.so is not built: run __graft_entry__.build()")
    return host_check.load()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_kernel_code_matches_reference(harness, oracle, case):
    field, planes, lut, opts, max_iter = kfp_planes_io.inputs(case, oracle)
    rgb = kfp_planes_io.colourise_host_check(harness, field, planes, lut, opts, max_iter)
    np.testing.assert_array_equal(rgb, load(case["name"]))


def test_planes_change_the_colours(harness, oracle):
    """The orbit planes are read: the same options without planes colour
    differently (the smoothing comes from test1, not from the field)."""
    case = CASES[0]
    field, planes, lut, opts, max_iter = kfp_planes_io.inputs(case, oracle)
    plain = host_check.colourise(host_check.bind_kfp(harness), field, max_iter, opts, lut)
    assert (plain != load(case["name"])).any()


def test_hidden_glitches_are_black(oracle):
    case = next(c for c in CASES if "transition_glitches" in c)
    _, planes, _, _, _ = kfp_planes_io.inputs(case, oracle)
    hidden = planes.arrays["transition"] < 0
    rgb = load(case["name"])
    assert hidden.sum() == case["transition_glitches"]
    assert (rgb[hidden] == 0).all()


@pytest.fixture(scope="module")
def hip(product):
    return native.bind_kfp(product)


def test_hip_rejects_bad_planes(hip, oracle):
    field, planes, lut, opts, max_iter = kfp_planes_io.inputs(CASES[0], oracle)
    planes.struct.version = 2
    with pytest.raises(RuntimeError, match="invalid native KFP colour"):
        kfp_planes_io.colourise(hip, field, planes, lut, opts, max_iter)
    planes.struct.version = 1
    planes.struct.texture_width = 5  # a texture size without texture bytes
    with pytest.raises(RuntimeError, match="invalid native KFP colour"):
        kfp_planes_io.colourise(hip, field, planes, lut, opts, max_iter)


def test_hip_planes_without_device_fail_loudly(hip, oracle):
    if hip.fractal_backend_capabilities() & 8:
        pytest.skip("a device is present")
    field, planes, lut, opts, max_iter = kfp_planes_io.inputs(CASES[0], oracle)
    with pytest.raises(RuntimeError):
        kfp_planes_io.colourise(hip, field, planes, lut, opts, max_iter)


def test_hip_atlas_planes_argument_checks(hip, oracle):
    """The plane-aware atlas entry's checks (renderer.cpp:17351-17388) run
    before any device work."""
    from kfp_planes_cases import ATLAS_CASES
    case = ATLAS_CASES[0]
    parent, child, lut, opts = kfp_planes_io.atlas_inputs(case, oracle)
    with pytest.raises(RuntimeError, match="invalid native plane-aware KFP atlas dimensions or controls"):
        kfp_planes_io.atlas_colourise(hip, dict(case, fraction=1.5), parent, child, lut, opts)
    with pytest.raises(RuntimeError, match="child absence"):
        kfp_planes_io.atlas_colourise(hip, dict(case, fraction=0.0), parent, child, lut, opts)
    with pytest.raises(RuntimeError, match="field bias exceeds"):
        kfp_planes_io.atlas_colourise(hip, case, parent, child, lut, native.KfpOptions.make(field_bias=30000.0))


@pytest.mark.parametrize("name", ["kfp_crop_default", "kfp_crop_stretch", "kfp_crop_centred"])
def test_crop_equals_the_childless_raw_atlas_oracle(oracle, name):
    """The KFP crop is the raw atlas pass without a child when the KFP field
    bias is 0 or the tile's cap (renderer.cpp:17803-17968 against
    17969-18265): the C oracle's raw atlas reproduces the reference's crop."""
    from kfp_planes_cases import CROP_CASES
    case = next(c for c in CROP_CASES if c["name"] == name)
    field, lut, opts = kfp_planes_io.crop_inputs(case, oracle)
    fn = oracle.fo_atlas_colourise_kfp_raw
    fn.argtypes = [native.c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, native.c_float_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, native.c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(native.KfpOptions),
                   native.c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    out = np.zeros((case["oh"], case["ow"], 3), dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    assert fn(field.ctypes.data_as(native.c_float_p), field.shape[1], field.shape[0], case["max_iter"], None, 0, 0, 0,
              out.ctypes.data_as(native.c_uint8_p), case["ow"], case["oh"], case["zoom"], 0.0, 1.0, case["max_iter"],
              ctypes.byref(opts), lut.ctypes.data_as(native.c_uint8_p), lut.shape[0], 0, err, 256) == 0, err.value
    np.testing.assert_array_equal(out, load(name))


from kfp_planes_cases import RECT_CASES  # noqa: E402


@pytest.mark.parametrize("case", RECT_CASES, ids=[c["name"] for c in RECT_CASES])
def test_rect_atlas_kernel_code_matches_reference(harness, oracle, case):
    """fractal_atlas_colourise_kfp's per-pixel code (kfp_core.h
    rect_atlas_pixel) on the CPU: the reference's bytes."""
    parent, child, lut, opts = kfp_planes_io.rect_inputs(case, oracle)
    rgb = kfp_planes_io.rect_host_check(harness, case, parent, child, lut, opts)
    np.testing.assert_array_equal(rgb, load(case["name"]))


def test_hip_rect_atlas_argument_checks(hip, oracle):
    case = RECT_CASES[0]
    parent, child, lut, opts = kfp_planes_io.rect_inputs(case, oracle)
    with pytest.raises(RuntimeError, match="child rectangle"):
        kfp_planes_io.rect_colourise(hip, dict(case, left=60), parent, child, lut, opts)
    with pytest.raises(RuntimeError, match="child absence"):
        kfp_planes_io.rect_colourise(hip, dict(case, feather=3), parent, None, lut, opts)
