"""Golden-vector cases for the deep-zoom field path.

Each case names the inputs of one reference creation plus one render call
through the renderer.h ABI (renderer.h:226-285).  ``make_golden.py`` runs
them through the reference's own renderer.cpp (compiled from source by
oracle/Makefile, with -ffp-contract=off) and stores the outputs next to this
file; the tests replay the same inputs through the C oracle and the HIP
library and compare.
"""
import random

BASE_X = "-0.743643887037158704752191506114774"
BASE_Y = "0.131825904205311970493132056385139"


def extended_centre(seed: int = 12345, digits: int = 100):
    """Base centre extended by ``digits`` seeded decimal digits per axis
    (the 1920x1080 zoom-1e100 workload's centre)."""
    rng = random.Random(seed)
    x = BASE_X + "".join(rng.choice("0123456789") for _ in range(digits))
    y = BASE_Y + "".join(rng.choice("0123456789") for _ in range(digits))
    return x, y


X100, Y100 = extended_centre()


def c3_centre(formula, zoom_exponent):
    """A near-boundary centre of c3_centres.json (make_c3_centres.py)."""
    import json
    from pathlib import Path
    table = json.loads((Path(__file__).resolve().parent / "c3_centres.json").read_text())
    entry = table[f"{formula}_{zoom_exponent}"]
    return entry["x"], entry["y"]


C3_30 = c3_centre(0, 30)
C3_100 = c3_centre(0, 100)

# The principal Misiurewicz point of the 1/3 limb plus a seeded 1e-101
# offset (fav_workloads.misiurewicz_field): every pixel escapes near
# iteration 810, and with the Kalles bailout some linear blocks end with
# |Z+d|^2 >= 3, so the endpoint replay runs.
MISIUREWICZ_X = ("-0.101096363845622161025785445738622565463805442826253483876931177660780840740470584274821219810"
                 "516779032980624190497440698")
MISIUREWICZ_Y = ("0.9562865108091415007710960577299774358098333365105291700343143215005246590657167325269784107873"
                 "39807194353461420359511709")


def _case(name, x, y, ref_zoom, zoom, max_iter, bits, width, height, series_order=8,
          series_block=32, image_series_order=8, clone_zoom=None, **options):
    return dict(name=name, x=x, y=y, ref_zoom=ref_zoom, zoom=zoom, max_iter=max_iter,
                bits=bits, width=width, height=height, series_order=series_order,
                series_block=series_block, image_series_order=image_series_order,
                clone_zoom=clone_zoom, options=options)


CASES = [
    # 640x360-workload geometry at reduced size: image series on, every pixel
    # finishes in the plain-double tail (render_fractal_ex parameters 8/32).
    _case("c1_series", BASE_X, BASE_Y, "1e12", "1e12", 20000, 128, 96, 54),
    _case("c1_noseries", BASE_X, BASE_Y, "1e12", "1e12", 20000, 128, 96, 54, series_min_terms=32),
    _case("c1_kalles", BASE_X, BASE_Y, "1e12", "1e12", 20000, 128, 80, 45,
          escape_radius_mode=1, coordinate_mode=1),
    # cubic + linear BLA blocks, double tail with rebases, escapes
    _case("bla_1e25", BASE_X, BASE_Y, "1e25", "1e25", 50000, 256, 64, 36, series_block=64,
          series_min_terms=32),
    _case("bla_1e27", BASE_X, BASE_Y, "1e27", "1e27", 50000, 256, 48, 27, series_block=64,
          series_min_terms=32),
    _case("bla_1e16_series", BASE_X, BASE_Y, "1e16", "1e16", 30000, 192, 48, 27, series_block=64),
    # 1e100 workload geometry (Kalles bailout/coordinates, series off,
    # order 3, block 256): all-interior, ~1563 linear jumps per pixel
    _case("c2_kalles", X100, Y100, "1e100", "1e100", 100000, 360, 48, 27, series_order=3,
          series_block=256, series_min_terms=32, escape_radius_mode=1, coordinate_mode=1),
    # BLA off at 1e100: every pixel hits the strict glitch sentinel (NaN)
    _case("glitch_1e100", X100, Y100, "1e100", "1e100", 100000, 360, 24, 14, series_order=3,
          series_block=256, series_min_terms=32, disable_bla=1),
    # the same with allow_recovery: glitches rebase instead of NaN
    _case("recovery_1e100", X100, Y100, "1e100", "1e100", 100000, 360, 24, 14, series_order=3,
          series_block=256, series_min_terms=32, disable_bla=1, allow_recovery=1),
    # non-strict render: Brent cycle detection terminates interior pixels
    _case("cycle_1e100", X100, Y100, "1e100", "1e100", 100000, 360, 32, 18, series_order=3,
          series_block=256, series_min_terms=32, strict=0),
    # deep and ultra-deep linear tiers: reference prepared for a wider view
    _case("deep_tier", X100, Y100, "1e60", "1e81", 100000, 360, 32, 18, series_order=3,
          series_block=4096, series_min_terms=32),
    _case("ultra_tier", X100, Y100, "1e60", "1e101", 100000, 360, 32, 18, series_order=3,
          series_block=4096, series_min_terms=32),
    # radius tier clone of a reusable root (renderer.cpp:13704-13768)
    _case("clone_tier", X100, Y100, "1e90", "1e100", 100000, 360, 32, 18, series_order=3,
          series_block=256, series_min_terms=32, clone_zoom="1e100"),
    # near-boundary precision centres (a period-121 / period-406 nucleus plus
    # a 1e-32 / 1e-102 offset), the 2560x1440 fields' views at 64x36: the
    # orbit returns near 0 every period, so the glitch test fires; strict
    # rendering leaves those pixels NaN, glitch rebase re-anchors them
    _case("c3_mandelbrot_1e30", *C3_30, "1e30", "1e30", 100000, 192, 64, 36, series_order=3, series_block=256,
          allow_recovery=1),
    _case("c3_mandelbrot_1e30_strict", *C3_30, "1e30", "1e30", 100000, 192, 64, 36, series_order=3,
          series_block=256),
    _case("c3_mandelbrot_1e100", *C3_100, "1e100", "1e100", 100000, 416, 64, 36, series_order=3,
          series_block=256, allow_recovery=1),
    _case("c3_mandelbrot_1e100_strict", *C3_100, "1e100", "1e100", 100000, 416, 32, 18, series_order=3,
          series_block=256),
    # BLA endpoint replay (renderer.cpp:13363-13514): linear blocks up to
    # 4096 steps, series off, Kalles bailout and grid
    _case("replay_misiurewicz", MISIUREWICZ_X, MISIUREWICZ_Y, "1e100", "1e100", 100000, 416, 64, 36, series_order=1,
          series_block=4096, series_min_terms=32, escape_radius_mode=1, coordinate_mode=1, allow_recovery=1),
]

STAT_NAMES = [
    "pixels", "logical_iterations", "bla_blocks", "linear_blocks", "cubic_blocks",
    "exact_steps", "replay_steps", "bla_retries", "cycle_inside", "double_tail_pixels",
    "bla_disabled_pixels", "tail_steps", "max_tail_steps", "tail_rebases",
    "tail_rebase_fallbacks", "max_pixel_iterations", "series_pixels", "series_jumps",
    "glitch_count", "unresolved_pixels", "deadline_aborts", "secondary_references",
    "render_ns",
] + [f"bla_length_{2 ** i}" for i in range(16)]
