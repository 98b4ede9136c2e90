"""Golden-vector cases for point renders (fractal_render_points,
renderer.h:287-321, renderer.cpp:20019-20186).

Each case names one reference creation and one point render.  The points
are drawn from a seeded generator (``make_points``) as mantissa pairs with
a shared exponent, deliberately not normalised (each mantissa pair is
scaled by a random power of two and the exponent compensates), plus special
values: dc = 0, a zero mantissa with a large exponent, negative zeros, and
points on the axes.  ``make_golden_points.py`` renders them through the
reference's own renderer.cpp on its scalar backend (backend 0) and stores
inputs and outputs next to this file.

Strict renders of at most 8192 points are chosen without glitches: there
the scalar backend repairs a glitched point with an exact MPFR iteration
(renderer.cpp:12944-12966), which the GPU point path does not do (it
returns NaN, renderer.cpp:996-1010).  The generator checks the glitch
counter of those cases.
"""
import math

import numpy as np

from alt_cases import BS_X, BS_Y, JULIA_C, JULIA_X, JULIA_Y, TC_X, TC_Y
from cases import BASE_X, BASE_Y, X100, Y100


def _case(name, x, y, ref_zoom, zoom, max_iter, bits, count, spread, seed, series_order=8,
          series_block=32, image_series_order=8, specials=True, fixed=None, formula=0, julia=("0", "0"),
          far=None, **options):
    return dict(name=name, x=x, y=y, ref_zoom=ref_zoom, zoom=zoom, max_iter=max_iter, bits=bits,
                count=count, spread=spread, seed=seed, series_order=series_order,
                series_block=series_block, image_series_order=image_series_order,
                specials=specials, fixed=fixed, formula=formula, julia=list(julia), far=far, options=options)


def half_height(zoom_text: str) -> float:
    """Half the project-coordinate view height, 1.4 / zoom (renderer.cpp:12692-12694)."""
    return 1.4 / float(zoom_text)


def make_points(case):
    """(real mantissas, imaginary mantissas, exponents) of a case."""
    if case["fixed"] is not None:
        re, im, e = case["fixed"]
        return (np.asarray(re, dtype=np.float64), np.asarray(im, dtype=np.float64),
                np.asarray(e, dtype=np.int32))
    rng = np.random.default_rng(case["seed"])
    n = case["count"]
    radius = half_height(case["zoom"]) * case["spread"]
    base = math.frexp(radius)[1]  # radius = m * 2^base, m in [0.5, 1)
    u = rng.uniform(-1.0, 1.0, size=(n, 2)) * radius
    shift = rng.integers(-3, 4, size=n)  # unnormalised mantissas, exact power-of-two scaling
    re = np.ldexp(u[:, 0], -(base - shift))
    im = np.ldexp(u[:, 1], -(base - shift))
    e = (base - shift).astype(np.int32)
    if case["specials"] and n >= 8:
        re[0], im[0], e[0] = 0.0, 0.0, 0                # dc = 0: the reference orbit itself
        re[1], im[1], e[1] = 0.0, 0.0, -4000            # zero mantissa, far exponent
        re[2], im[2] = -0.0, -0.0                       # negative zeros
        # the rest stay inside the radius of the random points
        re[3], im[3], e[3] = 0.0, 0.75, base - 1        # imaginary axis
        re[4], im[4], e[4] = -0.75, 0.0, base - 1       # real axis
        re[5], im[5], e[5] = 3.0, -5.0, base - 4        # mantissas well above one
        re[6], im[6], e[6] = 0.015625, 0.03125, base    # mantissas well below one half
    if case.get("far"):  # one more point at `far` half-heights: it alone sets the radius
        far_base = math.frexp(half_height(case["zoom"]) * case["far"])[1]
        re, im = np.append(re, 0.6), np.append(im, -0.6)
        e = np.append(e, np.int32(far_base))
    return re, im, e


CASES = [
    # the reference's own scalar-vs-GPU point case (tests/test_native_abi.py:1009-1060):
    # points 40x outside the view, so BLA is off by radius
    _case("points_reference_1e80", "-0.743643887037151", "0.13182590420533", "1e80", "1e80", 900, 768,
          12, 1.0, 0, series_order=3, series_block=64, specials=False,
          fixed=([0.0, 0.5, -0.5] * 4, [0.0, 0.25, -0.25] * 4, [-260] * 12),
          disable_bla=1, disable_cycle=1),
    # 1e12 with the image series on: points inside the view, escapes and interior
    _case("points_1e12_series", BASE_X, BASE_Y, "1e12", "1e12", 20000, 128, 300, 1.0, 11),
    # the same reference with points out to 3x the view: the radius exceeds the BLA
    # input radius, so BLA and the series are off (renderer.cpp:12701-12758);
    # non-strict, so glitched points stay NaN and the cycle check is on
    _case("points_1e12_far", BASE_X, BASE_Y, "1e12", "1e12", 20000, 128, 200, 3.0, 12, strict=0),
    # cubic + linear BLA blocks with escapes; glitched points rebase (recovery)
    _case("points_1e25_bla", BASE_X, BASE_Y, "1e25", "1e25", 50000, 256, 200, 1.0, 13,
          series_block=64, series_min_terms=32, strict=0, allow_recovery=1),
    # 1e100 workload geometry in project coordinates: ~1566 linear jumps per point
    _case("points_1e100", X100, Y100, "1e100", "1e100", 100000, 360, 128, 1.0, 14, series_order=3,
          series_block=256, series_min_terms=32, escape_radius_mode=1),
    # ultra-deep linear tier: reference prepared at 1e60, points of a 1e101 view
    _case("points_ultra_tier", X100, Y100, "1e60", "1e101", 100000, 360, 64, 1.0, 15, series_order=3,
          series_block=4096, series_min_terms=32),
    # BLA off at 1e100 with recovery: glitches rebase (non-strict, cycle check off)
    _case("points_1e100_recovery", X100, Y100, "1e100", "1e100", 100000, 360, 96, 1.0, 16,
          series_order=3, series_block=256, series_min_terms=32, disable_bla=1, allow_recovery=1,
          strict=0, disable_cycle=1),
    # BLA off at 1e100, strict, more points than the scalar backend repairs
    # exactly (8192): every glitched point is NaN on both paths
    _case("points_1e100_glitch_many", X100, Y100, "1e100", "1e100", 100000, 360, 8200, 1.0, 17,
          series_order=3, series_block=256, series_min_terms=32, disable_bla=1),
    # alternate formulas (render_alternate_reference_impl with points,
    # renderer.cpp:12064-12160): burning ship and tricorn BLA, the tricorn
    # cycle check, Julia with the image series on (points well inside its
    # radius) and off (the farthest point outside it)
    _case("points_burning_ship_1e7", BS_X, BS_Y, "1e7", "1e7", 4000, 200, 300, 1.0, 31, formula=2),
    _case("points_tricorn_cycle", TC_X, TC_Y, "2e7", "2e7", 5000, 200, 300, 1.0, 32, formula=3, strict=0),
    _case("points_julia_series", JULIA_X, JULIA_Y, "1e8", "1e8", 4000, 200, 300, 0.02, 33, formula=1,
          julia=JULIA_C),
    _case("points_julia_wide", JULIA_X, JULIA_Y, "1e8", "1e8", 4000, 200, 300, 1.0, 34, formula=1,
          julia=JULIA_C),
    # the same points and one far point: its radius turns the series off
    # (renderer.cpp:12121-12128), which changes some of the other values
    _case("points_julia_far", JULIA_X, JULIA_Y, "1e8", "1e8", 4000, 200, 300, 1.0, 34, formula=1,
          julia=JULIA_C, far=4.0),
]
