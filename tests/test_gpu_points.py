"""HIP point renders (fractal_render_points on backend 2) against the
fixtures the reference's renderer.cpp wrote on its scalar backend and
against the C oracle: NaN mask and counters identical, the float32 field
within the one-ulp allowance of the final long-double logarithm
(renderer.cpp:8170-8181)."""
import numpy as np
import pytest

import golden_io
import point_io
from point_cases import make_points

pytestmark = pytest.mark.gpu


def assert_points_parity(field, stats, expected_field, expected_stats):
    np.testing.assert_array_equal(np.isnan(field), np.isnan(expected_field))
    finite = ~np.isnan(expected_field)
    ulps = np.abs(field[finite].view(np.int32).astype(np.int64)
                  - expected_field[finite].view(np.int32).astype(np.int64))
    assert ulps.size == 0 or int(ulps.max()) <= golden_io.FIELD_ULPS, f"field differs by {int(ulps.max())} ulp"
    np.testing.assert_array_equal(golden_io.stats_without_time(stats), golden_io.stats_without_time(expected_stats))


@pytest.mark.parametrize("name", point_io.case_names())
def test_hip_points_match_reference(gpu_product, name):
    print(f"points case {name}", flush=True)
    golden = point_io.load(name)
    case = point_io.case_by_name(name)
    field, stats = point_io.render_abi(gpu_product, case, (golden["real"], golden["imag"], golden["exponent"]))
    assert_points_parity(field, stats, golden["field"], golden["stats"])


def test_hip_points_normalisation_is_exact(gpu_product):
    """Mantissas scaled by 2^k with the exponent lowered by k: the same field."""
    case = point_io.case_by_name("points_1e25_bla")
    golden = point_io.load(case["name"])
    k = np.random.default_rng(9).integers(-60, 61, size=golden["real"].size)
    points = (np.ldexp(golden["real"], k), np.ldexp(golden["imag"], k), (golden["exponent"] - k).astype(np.int32))
    field, stats = point_io.render_abi(gpu_product, case, points)
    assert_points_parity(field, stats, golden["field"], golden["stats"])


def test_hip_points_many_rows_match_oracle(gpu_product, oracle):
    """5000 points (19 full rows of 256 and a partial one) against the oracle,
    then a smaller render on the same handle (buffers reused)."""
    case = dict(point_io.case_by_name("points_1e12_series"), count=5000, seed=21)
    points = make_points(case)
    expected, expected_stats = point_io.render_oracle(oracle, case, points)
    field, stats = point_io.render_abi(gpu_product, case, points)
    assert_points_parity(field, stats, expected, expected_stats)
    small = tuple(a[:77] for a in points)
    expected, expected_stats = point_io.render_oracle(oracle, case, small)
    field, stats = point_io.render_abi(gpu_product, case, small)
    assert_points_parity(field, stats, expected, expected_stats)


def test_hip_single_point(gpu_product, oracle):
    case = point_io.case_by_name("points_1e100")
    golden = point_io.load(case["name"])
    one = (golden["real"][7:8], golden["imag"][7:8], golden["exponent"][7:8])
    expected, expected_stats = point_io.render_oracle(oracle, case, one)
    field, stats = point_io.render_abi(gpu_product, case, one)
    assert_points_parity(field, stats, expected, expected_stats)


@pytest.mark.parametrize("name", ["points_burning_ship_1e7", "points_julia_wide"])
def test_hip_alternate_points_many_rows_match_oracle(gpu_product, oracle, name):
    """Alternate-formula point renders of 3000 points (11 full rows and a
    partial one) against the oracle."""
    case = dict(point_io.case_by_name(name), count=3000, seed=41)
    points = make_points(case)
    expected, expected_stats = point_io.render_oracle(oracle, case, points)
    field, stats = point_io.render_abi(gpu_product, case, points)
    assert_points_parity(field, stats, expected, expected_stats)
