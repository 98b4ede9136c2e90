"""The C oracle's point render (fo_render_points) against the fixtures the
reference's own renderer.cpp wrote on its scalar backend
(fractal_render_points, renderer.cpp:20019-20186): field and work counters
identical to the bit, plus the exactness of the mantissa/exponent
normalisation the point ABI promises."""
import numpy as np
import pytest

import golden_io
import point_io
from point_cases import make_points


@pytest.mark.parametrize("name", point_io.case_names())
def test_fixture_inputs_regenerate(name):
    golden = point_io.load(name)
    re, im, e = make_points(point_io.case_by_name(name))
    np.testing.assert_array_equal(re.view(np.uint64), golden["real"].view(np.uint64))
    np.testing.assert_array_equal(im.view(np.uint64), golden["imag"].view(np.uint64))
    np.testing.assert_array_equal(e, golden["exponent"])


@pytest.mark.parametrize("name", point_io.case_names())
def test_oracle_points_match_reference(oracle, name):
    golden = point_io.load(name)
    case = point_io.case_by_name(name)
    field, stats = point_io.render_oracle(oracle, case, (golden["real"], golden["imag"], golden["exponent"]))
    np.testing.assert_array_equal(field.view(np.uint32), golden["field"].view(np.uint32))
    np.testing.assert_array_equal(golden_io.stats_without_time(stats), golden_io.stats_without_time(golden["stats"]))


def test_fixtures_cover_the_point_branches():
    """Escapes, interior, NaN glitches, rebases, BLA blocks and the radius
    cut-off all occur somewhere in the fixtures."""
    fields = {n: point_io.load(n) for n in point_io.case_names()}
    assert any(np.isnan(g["field"]).any() for g in fields.values())
    escaped = [g for g in fields.values() if (g["field"][np.isfinite(g["field"])] < g["meta"]["max_iter"]).any()]
    assert escaped
    assert fields["points_1e100"]["stats"][3] > 0               # linear BLA blocks
    assert fields["points_1e12_far"]["stats"][2] == 0           # radius beyond the BLA input radius
    assert fields["points_1e100_recovery"]["stats"][18] > 0     # glitches that rebased ...
    assert not np.isnan(fields["points_1e100_recovery"]["field"]).any()
    assert np.isnan(fields["points_1e100_glitch_many"]["field"]).all()  # ... and strict NaN


def test_normalisation_is_exact(oracle):
    """A point's mantissas scaled by 2^k with the exponent lowered by k is the
    same offset: the render does not change (sc_normalize is exact)."""
    case = point_io.case_by_name("points_1e12_series")
    golden = point_io.load(case["name"])
    rng = np.random.default_rng(5)
    k = rng.integers(-40, 41, size=golden["real"].size)
    points = (np.ldexp(golden["real"], k), np.ldexp(golden["imag"], k),
              (golden["exponent"] - k).astype(np.int32))
    field, _ = point_io.render_oracle(oracle, case, points, stats=False)
    np.testing.assert_array_equal(field.view(np.uint32), golden["field"].view(np.uint32))


def test_dropped_radius_is_caught(oracle):
    """The farthest point decides the radius of the series and BLA checks
    (renderer.cpp:12701-12758): the near points of the 3x-view case alone
    take the image series, the whole set does not, so a render that used
    the viewport's radius instead would differ in the counters compared
    above."""
    case = point_io.case_by_name("points_1e12_far")
    golden = point_io.load(case["name"])
    re, im, e = golden["real"], golden["imag"], golden["exponent"]
    magnitude = np.hypot(np.ldexp(re, e), np.ldexp(im, e))
    near = magnitude < 1.0e-12
    assert 0 < near.sum() < near.size
    _, stats = point_io.render_oracle(oracle, case, (re[near], im[near], e[near]))
    assert stats[16] == near.sum() and golden["stats"][16] == 0  # series pixels


def test_far_point_turns_the_julia_series_off():
    """The Julia series is taken only when the farthest point lies inside its
    radius (renderer.cpp:12121-12128): one far point changes values of the
    other points, so the fixtures pin that check."""
    wide = point_io.load("points_julia_wide")["field"]
    far = point_io.load("points_julia_far")["field"]
    assert far.size == wide.size + 1
    assert (wide.view(np.uint32) != far[:-1].view(np.uint32)).sum() > 0
