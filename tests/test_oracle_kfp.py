"""KFP colour transfer: the C oracle against golden RGB written by the
reference's renderer.cpp, and the palette LUT against its closed form
(the reference's own tests TVZ:1353-1405)."""
import json
import math

import numpy as np
import pytest

import native
import golden_io
from kfp_cases import CASES
from golden_io import load_kfp

GOLDEN = native.ROOT / "tests" / "golden"


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_oracle_kfp_matches_reference_golden(oracle, name):
    g = load_kfp(name)
    rgb = native.colourise_oracle(oracle, g["field"], g["max_iter"], g["options"], g["lut"])
    np.testing.assert_array_equal(rgb, g["rgb"])


def test_palette_lut_closed_form(oracle):
    """Kalles' cyclic sine interpolation, truncated (TVZ:1385-1405)."""
    stops = [(7, 19, 241), (240, 31, 11), (63, 222, 95)]
    size = 37
    lut = native.palette_lut_oracle(oracle, stops, size)
    expected = []
    for index in range(size):
        position = index * len(stops) / size
        left = int(math.floor(position))
        right = (left + 1) % len(stops)
        fraction = position - left
        fraction = math.sin((fraction - 0.5) * math.pi) / 2.0 + 0.5
        expected.append(tuple(int(fraction * stops[right][c] + (1.0 - fraction) * stops[left][c]) for c in range(3)))
    np.testing.assert_array_equal(lut, np.asarray(expected, dtype=np.uint8))


def test_palette_lut_wraps_from_last_stop_to_first(oracle):
    """TVZ:1353-1359: the table is cyclic, the last entry blends back."""
    lut = native.palette_lut_oracle(oracle, [(0, 0, 0), (255, 255, 255)], 16)
    assert tuple(lut[0]) == (0, 0, 0)
    assert int(lut[-1, 0]) < 255
    # at the half-way position between the two stops the sine blend is 1/2
    assert tuple(lut[4]) == (127, 127, 127) or tuple(lut[4]) == (128, 128, 128) or lut[4, 0] in (127, 128)


def test_dither_mask_closed_form():
    """Ordered mask ((x + 67c + 236y)*119 & 255)/256 (TVZ:1588-1612): a
    constant palette value v lands on floor(255 v + mask)."""
    field = np.full((3, 5), 10.0, dtype=np.float32)
    lut = np.full((1024, 3), 100, dtype=np.uint8)
    lib = native.load_oracle()
    opts = native.KfpOptions.make(color_method=0, slopes=0)
    rgb = native.colourise_oracle(lib, field, 1000, opts, lut)
    value = np.float32(100.0 / 255.0)
    for y in range(3):
        for x in range(5):
            for c in range(3):
                mask = np.float32(((x + 67 * c + 236 * y) * 119) & 255) / np.float32(256.0)
                assert rgb[y, x, c] == int(math.floor(np.float32(255.0) * value + mask))


def test_interior_pixels_take_the_interior_colour(oracle):
    field = np.full((4, 4), 5000.0, dtype=np.float32)
    field[0, 0] = 12.5
    lut = native.palette_lut_oracle(oracle, [(0, 0, 0), (255, 255, 255)])
    opts = native.KfpOptions.make(color_method=0, slopes=1, interior_color=(1, 2, 3))
    rgb = native.colourise_oracle(oracle, field, 5000, opts, lut)
    assert (rgb.reshape(-1, 3)[1:] == [1, 2, 3]).all()
    assert tuple(rgb[0, 0]) != (1, 2, 3)


@pytest.mark.parametrize("count,size,seed", [(2, 1024, 0), (8, 1024, 1), (12, 1024, 5), (3, 37, 2), (255, 1024, 9)])
def test_product_palette_lut_equals_oracle(oracle, count, size, seed):
    """The package's host-side LUT builder against the oracle's C table
    (both pinned to the closed form above)."""
    import fav_palette
    stops = fav_palette.seeded_stops(count, seed)
    np.testing.assert_array_equal(fav_palette.kfp_palette_lut(stops, size),
                                  native.palette_lut_oracle(oracle, stops, size))


def test_product_palette_lut_matches_golden_luts():
    """The LUTs stored with the reference's golden RGB were built by the same rule."""
    import fav_palette
    from kfp_cases import KALLES_DEFAULT_STOPS
    g = load_kfp("kalles_default_c1")
    np.testing.assert_array_equal(fav_palette.kfp_palette_lut(KALLES_DEFAULT_STOPS, 1024), g["lut"])
