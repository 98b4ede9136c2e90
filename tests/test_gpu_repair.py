"""The glitch-repair driver over the HIP library (secondary references and
fractal_render_points on the device) against the fixtures the reference's
renderer.cpp wrote through the same driver: the strict initial field has
the reference's NaN mask, and the repaired field equals the reference's
within the one-ulp allowance of the final long-double logarithm."""
import numpy as np
import pytest

import fav_repair
import golden_io
import native
import repair_io
from repair_cases import CASES

pytestmark = pytest.mark.gpu


def assert_within_ulp(field, expected):
    np.testing.assert_array_equal(np.isnan(field), np.isnan(expected))
    finite = ~np.isnan(expected)
    ulps = np.abs(field[finite].view(np.int32).astype(np.int64) - expected[finite].view(np.int32).astype(np.int64))
    assert ulps.size == 0 or int(ulps.max()) <= golden_io.FIELD_ULPS


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_hip_repair_matches_reference(gpu_product, case):
    print(f"repair case {case['name']}", flush=True)
    golden = repair_io.load(case["name"])
    zoom = fav_repair.zoom_text(case["log10_zoom"])
    bits = fav_repair.native_precision_bits(case["x"], case["y"], case["log10_zoom"])
    opts = native.RenderOptions.defaults(backend=2, **case["initial"])
    ref = native.AbiReference(gpu_product, case["x"], case["y"], zoom, case["max_iter"], bits,
                              repair_io.IMAGE_SERIES_ORDER, opts)
    initial, _, _ = ref.render(case["width"], case["height"], zoom, case["max_iter"], case["series_order"],
                               case["series_block"], opts)
    assert_within_ulp(initial, golden["initial"])
    diagnostics = {}
    repaired = fav_repair.repair_glitches(
        golden["initial"], log10_zoom=case["log10_zoom"], x_center=case["x"], y_center=case["y"],
        max_iter=case["max_iter"], series_order=case["series_order"], series_block=case["series_block"],
        diagnostics=diagnostics)
    assert_within_ulp(repaired, golden["repaired"])
    assert np.isfinite(repaired).all()
    assert diagnostics == golden["meta"]["diagnostics"]
