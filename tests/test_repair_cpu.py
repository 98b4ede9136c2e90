"""The glitch-repair driver (fav_repair.py, visualizer.py:8215-9252) on the
CPU: its geometry against exact rational arithmetic, its region search
against scipy's labelling, and the whole driver over the C oracle against
the fixtures the reference's renderer.cpp wrote through the same driver."""
from fractions import Fraction
from decimal import Decimal

import numpy as np
import pytest
from scipy import ndimage

import fav_repair
import repair_io
from repair_cases import CASES


def regions_by_labelling(mask):
    """Independent restatement: scipy's 8-connected labels, boxes with a
    one-pixel halo, rounded mean probes, largest first in raster order."""
    labels, count = ndimage.label(mask, structure=np.ones((3, 3), dtype=int))
    height, width = mask.shape
    out = []
    first = {}
    for index, label in enumerate(labels.ravel()):
        if label and label not in first:
            first[label] = index
    for label in sorted(first, key=first.get):
        ys, xs = np.nonzero(labels == label)
        out.append((max(0, int(xs.min()) - 1), min(width, int(xs.max()) + 2), max(0, int(ys.min()) - 1),
                    min(height, int(ys.max()) + 2), int(xs.size), int(round(xs.sum() / xs.size)),
                    int(round(ys.sum() / ys.size))))
    return sorted(out, key=lambda r: r[4], reverse=True)


@pytest.mark.parametrize("seed", range(6))
def test_regions_match_labelling(seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((23, 37)) < (0.05 + 0.1 * seed)
    assert fav_repair.unresolved_regions(mask) == regions_by_labelling(mask)


def test_regions_of_empty_and_full_masks():
    assert fav_repair.unresolved_regions(np.zeros((4, 5), dtype=bool)) == []
    assert fav_repair.unresolved_regions(np.ones((4, 5), dtype=bool)) == [(0, 5, 0, 4, 20, 2, 2)]
    with pytest.raises(ValueError):
        fav_repair.unresolved_regions(np.zeros(5, dtype=bool))


@pytest.mark.parametrize("coordinate_mode", [0, 1])
def test_local_reference_cell_is_the_probe_pixel_centre(coordinate_mode):
    """local centre = centre + view * (pixel offset) exactly (project pixel
    centres at (2p - (n-1)) / 2n, Kalles at (2p - n) / 2n), checked in exact
    rationals; the zoom fits the cell's larger ratio to the frame."""
    x, y, width, height = "-0.75", "0.1", 64, 36
    cell, probe = (10, 18, 4, 9), (13, 6)
    _, _, _, _, lx, ly, log_zoom = fav_repair.local_reference_cell(width, height, 30.0, x, y, cell, probe,
                                                                    coordinate_mode)
    view_h = Fraction(4 if coordinate_mode else 14, 1 if coordinate_mode else 5) / Fraction(10) ** 30
    view_w = view_h * width / height
    origin_x = width if coordinate_mode else width - 1
    origin_y = height if coordinate_mode else height - 1
    ex = Fraction(Decimal(x)) + view_w * Fraction(2 * probe[0] - origin_x, 2 * width)
    ey = Fraction(Decimal(y)) + view_h * Fraction(origin_y - 2 * probe[1], 2 * height)
    assert abs(Fraction(Decimal(lx)) - ex) < view_h * Fraction(1, 10 ** 40)
    assert abs(Fraction(Decimal(ly)) - ey) < view_h * Fraction(1, 10 ** 40)
    assert log_zoom == pytest.approx(30.0 + np.log10(max(64 / 8, 36 / 5)), abs=1e-12)


def test_point_offsets_are_pixel_steps_from_the_probe():
    width, height, cell, probe = 48, 27, (5, 17, 3, 11), (9, 6)
    re, im, e = fav_repair.point_offsets(width, height, 25.0, cell, probe, "-0.75", "0.1", "-0.75", "0.1")
    step = Fraction(14, 5) / Fraction(10) ** 25 / height  # view height / height = view width / width
    xs, ys = np.meshgrid(np.arange(5, 17), np.arange(3, 11))
    for r, i, k, px, py in zip(re, im, e, xs.ravel(), ys.ravel()):
        exact_re, exact_im = step * int(px - probe[0]), step * int(probe[1] - py)
        got_re, got_im = Fraction(float(r)) * Fraction(2) ** int(k), Fraction(float(i)) * Fraction(2) ** int(k)
        assert abs(got_re - exact_re) <= step * Fraction(1, 10 ** 12)
        assert abs(got_im - exact_im) <= step * Fraction(1, 10 ** 12)
        assert max(abs(r), abs(i)) == 0.0 or 0.5 <= max(abs(r), abs(i)) < 1.0  # shared exponent of the larger


@pytest.mark.parametrize("text", ["1.5e-300", "-7.25e40", "3", "0.1"])
def test_decimal_to_scaled(text):
    m, e = fav_repair.decimal_to_scaled(Decimal(text))
    assert 0.5 <= abs(m) < 1.0
    assert abs(Fraction(m) * Fraction(2) ** e / Fraction(Decimal(text)) - 1) < Fraction(1, 10 ** 13)
    assert fav_repair.decimal_to_scaled(Decimal(0)) == (0.0, 0)


def test_zoom_text():
    assert fav_repair.zoom_text(60.0) == "1e+60"
    assert float(fav_repair.zoom_text(12.5)) == pytest.approx(10 ** 12.5, rel=1e-15)


def test_clean_field_is_returned_unchanged():
    field = np.arange(12, dtype=np.float32).reshape(3, 4)
    out = fav_repair.repair_glitches(field, log10_zoom=20.0, x_center="-0.75", y_center="0.1", max_iter=100,
                                     reference_factory=lambda *a: pytest.fail("no reference is needed"))
    np.testing.assert_array_equal(out, field)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_repair_over_oracle_matches_reference(oracle, case):
    golden = repair_io.load(case["name"])
    diagnostics = {}
    repaired = fav_repair.repair_glitches(
        golden["initial"], log10_zoom=case["log10_zoom"], x_center=case["x"], y_center=case["y"],
        max_iter=case["max_iter"], series_order=case["series_order"], series_block=case["series_block"],
        reference_factory=lambda *a: repair_io.OracleSecondary(oracle, *a), diagnostics=diagnostics)
    np.testing.assert_array_equal(repaired.view(np.uint32), golden["repaired"].view(np.uint32))
    assert diagnostics == golden["meta"]["diagnostics"]
    assert np.isfinite(repaired).all() and np.isnan(golden["initial"]).any()


def load_visualizer_fixture(case):
    import json
    from pathlib import Path
    path = Path(__file__).resolve().parent / "golden" / f"repair_visualizer_{case['name'][len('repair_'):]}.npz"
    data = np.load(path, allow_pickle=False)
    return json.loads(str(data["meta"])), data["field"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_driver_against_the_reference_repair_function(case):
    """The reference's own `_atlas_glitch_reference_field` (run over its
    renderer.cpp by tests/golden/make_golden_repair_visualizer.py, from the
    same strict initial field) against this driver's fixture, which the
    driver reproduces over the oracle and the HIP path (tests above and
    tests/test_gpu_repair.py):

    * a region large enough for the region queue (the 9216-pixel all-NaN
      BLA-off field at 1e100, split into four secondary references): the
      whole repaired field bit for bit;
    * 14 scattered glitch pixels (Old Wood Dish at 1e60): the reference
      takes its exact arbitrary-precision pixel retry for masks this small
      (visualizer.py:8472-8526, 8640-8650), which this driver leaves out.
      Its values equal the driver's secondary-reference values bit for bit
      on the 14 pixels; around them the driver, as the reference's region
      path does (visualizer.py:8905-8906), also takes the secondary render's
      finite halo pixels, which the exact retry leaves at their initial
      values."""
    meta, reference_field = load_visualizer_fixture(case)
    golden = repair_io.load(case["name"])
    initial, repaired = golden["initial"], golden["repaired"]
    glitched = np.isnan(initial)
    if meta["diagnostics"]["exact_pixel_recoveries"] == 0:
        np.testing.assert_array_equal(reference_field.view(np.uint32), repaired.view(np.uint32))
        assert meta["diagnostics"]["secondary_references"] == golden["meta"]["diagnostics"]["secondary_references"]
        return
    assert meta["diagnostics"]["exact_pixel_recoveries"] == int(glitched.sum())
    np.testing.assert_array_equal(reference_field[glitched].view(np.uint32), repaired[glitched].view(np.uint32))
    np.testing.assert_array_equal(reference_field[~glitched].view(np.uint32), initial[~glitched].view(np.uint32))
    changed = repaired != initial
    boxes = np.zeros_like(glitched)
    for x0, x1, y0, y1, *_ in fav_repair.unresolved_regions(glitched):
        boxes[y0:y1, x0:x1] = True
    assert not (changed & ~glitched & ~boxes).any()  # the driver only touches the regions' boxes


def test_spatial_recovery_grows_means_inward():
    """A 3x3 hole in a linear ramp: the ring next to the finite pixels is
    the mean of its finite neighbours, the centre the mean of the ring;
    fallback values are taken first where finite."""
    field = np.add.outer(np.arange(7.0), 10.0 * np.arange(7.0)).astype(np.float32)
    hole = field.copy()
    hole[2:5, 2:5] = np.nan
    out, filled = fav_repair.spatial_recover_field(hole)
    assert filled == 9 and np.isfinite(out).all()
    np.testing.assert_array_equal(out[~np.isnan(hole)], field[~np.isnan(hole)])
    # corner of the hole: neighbours (1,1) (1,2) (1,3) (2,1) (3,1) are finite
    expect = np.mean([field[1, 1], field[1, 2], field[1, 3], field[2, 1], field[3, 1]])
    assert out[2, 2] == pytest.approx(expect, rel=1e-6)
    # the centre (the ramp is linear, so the symmetric ring averages to the exact value)
    assert out[3, 3] == pytest.approx(field[3, 3], rel=1e-6)
    fallback = np.full_like(field, np.nan)
    fallback[3, 3] = -5.0
    out, filled = fav_repair.spatial_recover_field(hole, fallback)
    assert out[3, 3] == -5.0 and filled == 9
    with pytest.raises(RuntimeError):
        fav_repair.spatial_recover_field(np.full((3, 3), np.nan, dtype=np.float32))
