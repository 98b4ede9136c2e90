"""Pins of the oracle that do not come from the reference's outputs.

Each sampled pixel is iterated directly, z <- z^2 + c, in arbitrary
precision (mpmath) from the pixel-centre coordinate that the viewport
mapping defines (renderer.cpp:142-158, 12689-12783).  A wrong axis sign,
origin, pixel spacing, escape radius, BLA coefficient or tail rebase moves
the escape iteration of these pixels.  The tolerances follow the
reference's own mpmath tests (TNA:1522-1592 <=0.02 at 1e12, TNA:1594-1664
<=0.03 at 1e30), and escape counts must agree exactly.
"""
import math

import mpmath
import numpy as np
import pytest

import golden_io


def pixel_c(meta, px, py):
    """Pixel centre from the project/Kalles viewport conventions."""
    w, h = meta["width"], meta["height"]
    kalles = meta["options"].get("coordinate_mode", 0) == 1
    zoom = mpmath.mpf(meta["zoom"])
    view_h = (4 if kalles else mpmath.mpf("2.8")) / zoom
    view_w = view_h * w / h
    ox = (w // 2) if kalles else mpmath.mpf(w - 1) / 2
    oy = (h // 2) if kalles else mpmath.mpf(h - 1) / 2
    x = mpmath.mpf(meta["x"]) + (px - ox) / mpmath.mpf(w) * view_w
    y = mpmath.mpf(meta["y"]) - (py - oy) / mpmath.mpf(h) * view_h
    return mpmath.mpc(x, y)


def direct_escape(c, max_iter, radius2):
    """(n, smooth) with z_1 = c; n is the first index with |z_n|^2 > R^2."""
    z = c
    for n in range(1, max_iter):
        m2 = z.real * z.real + z.imag * z.imag
        if m2 > radius2:
            return n, float(n - mpmath.log(mpmath.log(mpmath.sqrt(m2))) / mpmath.log(2))
        z = z * z + c
    return max_iter, float(max_iter)


@pytest.mark.parametrize("name,samples,tol", [
    ("c1_noseries", 24, 0.02),
    ("c1_series", 16, 0.02),
    ("c1_kalles", 16, 0.02),
    ("bla_1e25", 12, 0.03),
    ("bla_1e27", 10, 0.03),
])
def test_oracle_agrees_with_direct_iteration(oracle, name, samples, tol):
    golden = golden_io.load(name)
    meta = golden["meta"]
    field, planes, _ = golden_io.render_oracle(oracle, meta, stats=False)
    radius2 = 1e8 if meta["options"].get("escape_radius_mode", 0) == 1 else 4
    mpmath.mp.prec = meta["bits"] + 64
    rng = np.random.default_rng(7)
    w, h = meta["width"], meta["height"]
    for _ in range(samples):
        px, py = int(rng.integers(0, w)), int(rng.integers(0, h))
        n, smooth = direct_escape(pixel_c(meta, px, py), meta["max_iter"], radius2)
        idx = py * w + px
        count = int(planes.orbit_iteration[idx]) + 1
        assert count == n, (name, px, py, count, n)
        assert abs(float(field[py, px]) - smooth) <= tol, (name, px, py, field[py, px], smooth)


def test_interior_workload_is_bounded_under_direct_iteration(oracle):
    """The 1e100 workload's pixels stay bounded for the full cap."""
    golden = golden_io.load("c2_kalles")
    meta = golden["meta"]
    assert (golden["orbit_iteration"] == meta["max_iter"]).all()
    mpmath.mp.prec = meta["bits"] + 32
    c = pixel_c(meta, 0, 0)  # a corner, the pixel farthest from the centre
    z = c
    for _ in range(meta["max_iter"] // 10):
        z = z * z + c
        assert abs(z) <= 2
    field, _, _ = golden_io.render_oracle(oracle, meta, stats=False)
    assert np.all(field == np.float32(meta["max_iter"]))


@pytest.mark.parametrize("name", ["c1_noseries", "bla_1e25"])
def test_smooth_value_closed_form_from_planes(name):
    """field = n - log2(ln|z_n|) with |z_n|^2 = test1 (renderer.cpp:8170-8188,
    12578-12585): the FP64 smooth field derived from the planes."""
    golden = golden_io.load(name)
    escaped = golden["orbit_iteration"] < golden["meta"]["max_iter"]
    n = golden["orbit_iteration"][escaped].astype(np.float64) + 1
    smooth = n - np.log(0.5 * np.log(golden["test1"][escaped])) / math.log(2)
    np.testing.assert_allclose(golden["field"].reshape(-1)[escaped], smooth, rtol=0, atol=2e-3)
    assert np.all(golden["test1"][escaped] > (1e8 if golden["meta"]["options"].get("escape_radius_mode") else 4))
    assert np.all(golden["test2"][escaped] <= golden["test1"][escaped])
