"""fo_render_crop, the oracle's exact-geometry rectangle render, equals the
matching block of the full render (Mandelbrot with planes, and the burning
ship), so the GPU tests may compare large frames on bands."""
import numpy as np
import pytest

import golden_io
import native


@pytest.mark.parametrize("name", ["c1_series", "c3_mandelbrot_1e30_strict"])
def test_mandelbrot_crop_equals_full_render(oracle, name):
    m = golden_io.load(name)["meta"]
    opts = native.RenderOptions.defaults(**m["options"])
    ref = native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"],
                                 m["image_series_order"], m["options"].get("escape_radius_mode", 0),
                                 m["options"].get("coordinate_mode", 0))
    full, fp, _ = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"],
                             m["series_block"], opts, planes=True)
    for x, y, w, h in [(0, 0, m["width"], 1), (5, 3, 20, 11), (m["width"] - 1, m["height"] - 2, 1, 2)]:
        crop, cp, _ = ref.render_crop(m["width"], m["height"], (x, y, w, h), m["zoom"], m["max_iter"],
                                      m["series_order"], m["series_block"], opts, planes=True)
        np.testing.assert_array_equal(crop.view(np.uint32), full[y:y + h, x:x + w].view(np.uint32))
        for plane in native.PlaneArrays.NAMES:
            whole = getattr(fp, plane).reshape(m["height"], m["width"])[y:y + h, x:x + w].reshape(-1)
            np.testing.assert_array_equal(getattr(cp, plane), whole, err_msg=plane)


def test_alternate_crop_equals_full_render(oracle):
    m = golden_io.load("c3_burning_ship_1e30")["meta"]
    ref = native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], 8,
                                 formula=m["formula"])
    opts = native.RenderOptions.defaults(**m["options"])
    full, _, _ = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], 8, 32, opts)
    crop, _, stats = ref.render_crop(m["width"], m["height"], (7, 2, 30, 9), m["zoom"], m["max_iter"], 8, 32, opts,
                                     stats=True)
    np.testing.assert_array_equal(crop.view(np.uint32), full[2:11, 7:37].view(np.uint32))
    assert int(stats[0]) == 30 * 9


def test_crop_outside_the_field_is_rejected(oracle):
    m = golden_io.load("c1_series")["meta"]
    ref = native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], 8)
    with pytest.raises(RuntimeError, match="crop rectangle"):
        ref.render_crop(m["width"], m["height"], (m["width"] - 3, 0, 4, 1), m["zoom"], m["max_iter"])
