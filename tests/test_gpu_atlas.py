"""Raw KFP atlas colour pass on the GPU against the reference's golden
vectors and, at frame size, against the C oracle.

Bar: RGB within one 8-bit level (the device's log/pow/atan2/sin may round
differently from host libm at a quantisation step), as for the KFP pass."""
import numpy as np
import pytest

import atlas_io
import golden_io
from fav_workloads import atlas_tile as synthetic_tile

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", atlas_io.case_names())
def test_hip_atlas_matches_reference_golden(gpu_product, oracle, name):
    print(f"[gpu] atlas golden {name}", flush=True)
    g = atlas_io.load(name)
    parent, child, lut, opts = atlas_io.case_inputs(oracle, g["meta"])
    for entry in ("fractal_atlas_colourise_kfp_raw", "fractal_atlas_colourise_kfp_raw_precise"):
        fn = atlas_io.bind_atlas(gpu_product, entry)
        status, rgb = atlas_io.atlas_call(fn, g["meta"], parent, child, lut, opts)
        assert status == 0, gpu_product.fractal_last_error()
        golden_io.assert_rgb_within_one_level(rgb, g["rgb"])


def test_hip_atlas_frame_matches_oracle(gpu_product, oracle):
    """A 1920x1080 frame from 1280x720 parent and child tiles (child a third
    of the frame, zoomed 1.6x) against the C oracle's restatement."""
    print("[gpu] atlas 1920x1080 frame", flush=True)
    from test_atlas_cpu import oracle_call
    import ctypes
    import native
    oracle.fo_atlas_colourise_kfp_raw.argtypes = (
        atlas_io.ATLAS_ARGS[:15] + [ctypes.POINTER(native.KfpOptions), ctypes.POINTER(ctypes.c_uint8),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int])
    case = dict(atlas_io.CASES[0], output=[1920, 1080], parent=[1280, 720, 20000], child=[1280, 720, 24000],
                parent_zoom=1.35, child_fraction=0.33, child_zoom=1.6, max_iter=24000)
    parent = synthetic_tile(1280, 720, 20000, seed=31)
    child = synthetic_tile(1280, 720, 24000, seed=32)
    _, _, lut, opts = atlas_io.case_inputs(oracle, case)
    fn = atlas_io.bind_atlas(gpu_product)
    status, rgb = atlas_io.atlas_call(fn, case, parent, child, lut, opts)
    assert status == 0, gpu_product.fractal_last_error()
    expect = oracle_call(oracle, case, parent, child, lut, opts)
    golden_io.assert_rgb_within_one_level(rgb, expect)
    assert np.mean(rgb == expect) > 0.999
