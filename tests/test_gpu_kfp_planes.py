"""The HIP plane-aware KFP colour pass against the reference's RGB: every
channel within one 8-bit level (device log / pow / atan may round
differently from the host libm at a quantisation step), almost all exact."""
import numpy as np
import pytest

import golden_io
import kfp_planes_io
import native
from kfp_planes_cases import ATLAS_CASES, CASES, CROP_CASES
from test_kfp_planes_cpu import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_hip_kfp_planes_match_reference(gpu_product, oracle, case):
    print(f"[gpu] kfp planes {case['name']}", flush=True)
    field, planes, lut, opts, max_iter = kfp_planes_io.inputs(case, oracle)
    rgb = kfp_planes_io.colourise(native.bind_kfp(gpu_product), field, planes, lut, opts, max_iter)
    expected = load(case["name"])
    golden_io.assert_rgb_within_one_level(rgb, expected, case["name"])
    exact = float(np.mean(rgb == expected))
    print(f"[gpu] kfp planes {case['name']}: exact channel fraction {exact:.6f}", flush=True)
    assert exact > 0.995


@pytest.mark.parametrize("case", ATLAS_CASES, ids=[c["name"] for c in ATLAS_CASES])
@pytest.mark.parametrize("entry", ["fractal_atlas_colourise_kfp_planes", "fractal_atlas_colourise_kfp_planes_precise"])
def test_hip_atlas_kfp_planes_match_reference(gpu_product, oracle, case, entry):
    """Parent / child tiles and planes composed and coloured on the device."""
    print(f"[gpu] kfp atlas planes {case['name']} {entry}", flush=True)
    parent, child, lut, opts = kfp_planes_io.atlas_inputs(case, oracle)
    rgb = kfp_planes_io.atlas_colourise(gpu_product, case, parent, child, lut, opts, entry)
    expected = load(case["name"])
    golden_io.assert_rgb_within_one_level(rgb, expected, case["name"])
    exact = float(np.mean(rgb == expected))
    print(f"[gpu] kfp atlas planes {case['name']}: exact channel fraction {exact:.6f}", flush=True)
    assert exact > 0.995


@pytest.mark.parametrize("case", CROP_CASES, ids=[c["name"] for c in CROP_CASES])
@pytest.mark.parametrize("entry", ["fractal_crop_colourise_kfp", "fractal_crop_colourise_kfp_precise"])
def test_hip_crop_kfp_matches_reference(gpu_product, oracle, case, entry):
    print(f"[gpu] kfp crop {case['name']} {entry}", flush=True)
    field, lut, opts = kfp_planes_io.crop_inputs(case, oracle)
    rgb = kfp_planes_io.crop_colourise(gpu_product, case, field, lut, opts, entry)
    expected = load(case["name"])
    golden_io.assert_rgb_within_one_level(rgb, expected, case["name"])
    assert float(np.mean(rgb == expected)) > 0.995


from kfp_planes_cases import RECT_CASES  # noqa: E402


@pytest.mark.parametrize("case", RECT_CASES, ids=[c["name"] for c in RECT_CASES])
@pytest.mark.parametrize("entry", ["fractal_atlas_colourise_kfp", "fractal_atlas_colourise_kfp_precise"])
def test_hip_rect_atlas_matches_reference(gpu_product, oracle, case, entry):
    print(f"[gpu] kfp rect atlas {case['name']} {entry}", flush=True)
    parent, child, lut, opts = kfp_planes_io.rect_inputs(case, oracle)
    rgb = kfp_planes_io.rect_colourise(gpu_product, case, parent, child, lut, opts, entry)
    expected = load(case["name"])
    golden_io.assert_rgb_within_one_level(rgb, expected, case["name"])
    assert float(np.mean(rgb == expected)) > 0.995
