"""Shallow direct renderer: the C oracle against golden vectors written by
the reference's renderer.cpp (render_fractal_ex[_planes], scalar backend,
use_perturbation = 0), plus closed-form pins."""
import math

import numpy as np
import pytest

import native
from golden_io import load_direct
from direct_cases import CASES

def assert_direct_bit_exact(golden, field, planes):
    np.testing.assert_array_equal(field.reshape(-1).view(np.uint32), golden["field"].reshape(-1).view(np.uint32))
    if "orbit_iteration" in golden:
        for name in native.PlaneArrays.NAMES:
            np.testing.assert_array_equal(getattr(planes, name), golden[name], err_msg=name)


@pytest.mark.parametrize("name", [c["name"] for c in CASES])
def test_oracle_direct_matches_reference_golden(oracle, name):
    g = load_direct(name)
    field, planes = native.render_direct_oracle(oracle, g["meta"])
    assert_direct_bit_exact(g, field, planes)


def test_direct_escape_iteration_closed_form(oracle):
    """c = 1: z = 1, 2, 5 -> |z|^2 = 25 > 4 at the third sample, so the
    zero-based escape index is 2 and the field is 3 - log(log 5)/log 2."""
    case = dict(width=1, height=1, zoom="1", x="1", y="0", max_iter=100, formula=0, planes=True, options={})
    field, planes = native.render_direct_oracle(oracle, case)
    assert int(planes.orbit_iteration[0]) == 2
    assert planes.test1[0] == 25.0 and planes.test2[0] == 4.0
    assert field[0, 0] == np.float32(3.0 - math.log(math.log(5.0)) / math.log(2.0))


def test_direct_cardioid_and_bulb_are_interior(oracle):
    """c = 0 (cardioid) and c = -1 (period-2 bulb) never escape."""
    for x in ("0", "-1"):
        case = dict(width=1, height=1, zoom="1", x=x, y="0", max_iter=50, formula=0, planes=True, options={})
        field, planes = native.render_direct_oracle(oracle, case)
        assert field[0, 0] == 50.0 and int(planes.orbit_iteration[0]) == 50
        assert planes.test1[0] == 0.0 and planes.de_x[0] == 0.0


def test_direct_brute_force_tiny_fields(oracle):
    """Escape indices of a 12x7 field of each formula against a plain
    Python iteration of z -> f(z) + c."""
    maps = {
        0: lambda x, y, cx, cy: (x * x - y * y + cx, 2.0 * x * y + cy),
        2: lambda x, y, cx, cy: (abs(x) * abs(x) - abs(y) * abs(y) + cx, 2.0 * abs(x) * abs(y) + cy),
        3: lambda x, y, cx, cy: (x * x - y * y + cx, -2.0 * x * y + cy),
    }
    for formula, step in maps.items():
        case = dict(width=12, height=7, zoom="0.9", x="-0.4", y="0.1", max_iter=60, formula=formula, planes=True,
                    options={})
        _, planes = native.render_direct_oracle(oracle, case)
        hs = 2.8 / 0.9
        ws = hs * 12 / 7
        for py in range(7):
            for px in range(12):
                cx = -0.4 + (px - 5.5) * ws / 12
                cy = 0.1 + -(py - 3.0) * hs / 7
                if formula == 0:
                    q = (cx - 0.25) * (cx - 0.25) + cy * cy
                    if q * (q + cx - 0.25) <= 0.25 * cy * cy or (cx + 1.0) * (cx + 1.0) + cy * cy <= 0.0625:
                        assert planes.orbit_iteration[py * 12 + px] == 60
                        continue
                x = y = 0.0
                n = 60
                for it in range(60):
                    x, y = step(x, y, cx, cy)
                    if x * x + y * y > 4.0:
                        n = it
                        break
                assert planes.orbit_iteration[py * 12 + px] == n, (formula, px, py)


def test_direct_julia_derivative_starts_at_one(oracle):
    """Julia DE uses dz/dz0 with dz0 = 1: for c = 0 and z0 = 3 the first
    sample z = 9 escapes with dz = 6, so DE = |z| log|z| / (|dz| spacing)."""
    case = dict(width=1, height=1, zoom="1", x="3", y="0", max_iter=10, formula=1, julia_re=0.0, julia_im=0.0,
                planes=True, options={})
    _, planes = native.render_direct_oracle(oracle, case)
    spacing = 2.8 / 1.0
    expected = 9.0 * math.log(9.0) / (6.0 * spacing)
    assert planes.de_x[0] == pytest.approx(expected, rel=1e-12)
    assert planes.de_y[0] == 0.0
