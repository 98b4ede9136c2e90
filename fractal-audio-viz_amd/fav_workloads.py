"""Synthetic inputs of the measured deep-zoom workloads.

The centres are the reference's bundled deep point
(-0.743643887037158704752191506114774 + 0.131825904205311970493132056385139i)
optionally extended by seeded decimal digits, so the 1e100 field has a
well-defined 360-bit reference orbit.
"""
import random

BASE_X = "-0.743643887037158704752191506114774"
BASE_Y = "0.131825904205311970493132056385139"


def extended_centre(seed: int = 12345, digits: int = 100):
    rng = random.Random(seed)
    x = BASE_X + "".join(rng.choice("0123456789") for _ in range(digits))
    y = BASE_Y + "".join(rng.choice("0123456789") for _ in range(digits))
    return x, y


def field_1e100_1080p():
    """1920x1080 at zoom 1e100, 360-bit orbit, cap 100000, BLA on, rebase
    on, Kalles bailout (R=1e4) and pixel grid, image series off
    (series_min_terms=32 with an order-8 series, as the reference's
    benchmark.py:48-55 does), BLA order 3 with block 256, planes on (the
    distance/slope shading inputs)."""
    x, y = extended_centre()
    return dict(name="field_1920x1080_zoom1e100_cap100000", x=x, y=y, zoom="1e100", width=1920,
                height=1080, max_iter=100000, bits=360, image_series_order=8, series_order=3,
                series_block=256, options=dict(series_min_terms=32, escape_radius_mode=1,
                                               coordinate_mode=1), planes=True)


def field_1e12_640x360():
    """640x360 at zoom 1e12, 128-bit orbit, cap 20000, project grid,
    R=2, image series on (render_fractal_ex's order 8 / block 32)."""
    return dict(name="field_640x360_zoom1e12_cap20000", x=BASE_X, y=BASE_Y, zoom="1e12", width=640,
                height=360, max_iter=20000, bits=128, image_series_order=8, series_order=8,
                series_block=32, options={}, planes=True)


# Algorithmic FP64 operations per event (the roofline numerator).  Complex
# multiply 6, complex add 2, squared norm 3; normalisation, exponent and
# comparison work is not counted.  A linear block is A*d + B*dc (14) plus
# the |d|^2 that the BLA level lookup compares with the validity radii
# before every block (renderer.cpp:8816-8877), 3: the loop-top glitch,
# escape and endpoint tests of a block are properties of the reference
# index alone while d is negligible, so they are not per-pixel work.
FLOPS = {
    "exact_step": 16 + 16,    # d' = 2Zd + d^2 + dc, plus |Z+d|^2 margin and |d|^2
    "replay_step": 16 + 16,
    "tail_step": 19,          # plain-double tail step (renderer.cpp:12527-12554)
    "linear_block": 14 + 3,   # A*d + B*dc, plus |d|^2 for the level lookup
    "cubic_block": 70 + 16,   # order-3 Horner (renderer.cpp:8710-8730) with its margin
    "derivative": 10,         # dz' = 2 z dz + 1, per step when planes are written
}


def flops_from_stats(stats, planes: bool) -> float:
    """FP64 operations of one field from its RenderStats counters
    (renderer.cpp:4615-4643).  The derivative updates of linear blocks are
    left out: the field kernel defers them to a per-wave log and runs them
    only for pixels that go on to read the derivative (none on an interior
    field), so the count is a lower bound on the work done."""
    linear, cubic, exact, replay, tail = (float(stats[3]), float(stats[4]), float(stats[5]),
                                          float(stats[6]), float(stats[11]))
    scaled_exact = exact - tail  # exact_steps includes the plain-double tail steps
    total = (scaled_exact * FLOPS["exact_step"] + replay * FLOPS["replay_step"]
             + tail * FLOPS["tail_step"] + linear * FLOPS["linear_block"]
             + cubic * FLOPS["cubic_block"])
    if planes:
        total += (scaled_exact + replay + tail) * FLOPS["derivative"]
    return total


def julia_parameter(seed: int) -> tuple[str, str]:
    """Julia parameter c = e^{i theta}/2 - e^{2 i theta}/4 + 1e-3 e^{i phi},
    theta and phi drawn from random.Random(seed): a point 1e-3 from the main
    cardioid's boundary, as decimal text of the float64 values."""
    import math
    rng = random.Random(seed)
    theta, phi = rng.uniform(0.0, 2.0 * math.pi), rng.uniform(0.0, 2.0 * math.pi)
    re = math.cos(theta) / 2 - math.cos(2 * theta) / 4 + 1e-3 * math.cos(phi)
    im = math.sin(theta) / 2 - math.sin(2 * theta) / 4 + 1e-3 * math.sin(phi)
    return repr(re), repr(im)


def julia_fixed_point(c_re: str, c_im: str, digits: int = 384) -> tuple[str, str]:
    """The fixed point of z^2 + c with the larger modulus (the root of
    z^2 - z + c, (1 +- sqrt(1 - 4c)) / 2), as `digits`-digit decimal text:
    a point of the Julia set, whose orbit is constant."""
    from decimal import Decimal, localcontext
    with localcontext() as ctx:
        ctx.prec = digits + 20
        cr, ci = Decimal(c_re), Decimal(c_im)
        a, b = 1 - 4 * cr, -4 * ci
        modulus = (a * a + b * b).sqrt()
        sr = ((modulus + a) / 2).sqrt()
        si = ((modulus - a) / 2).sqrt()
        if b < 0:
            si = -si
        roots = [((1 + sr) / 2, si / 2), ((1 - sr) / 2, -si / 2)]
        zr, zi = max(roots, key=lambda z: z[0] * z[0] + z[1] * z[1])
        ctx.prec = digits
        return str(+zr), str(+zi)


def julia_field(seed: int = 1, width: int = 3840, height: int = 2160, zoom: str = "1e100",
                max_iter: int = 200000):
    """A Julia field viewed at its repelling fixed point (the reference's
    Julia view centre), 400-bit orbit, default render options."""
    c = julia_parameter(seed)
    x, y = julia_fixed_point(*c)
    return dict(name=f"julia_{width}x{height}_zoom{zoom}_cap{max_iter}_seed{seed}", formula=1, julia=c, x=x, y=y,
                zoom=zoom, width=width, height=height, max_iter=max_iter, bits=400)


# Tricorn preperiodic boundary roots of the reference's point catalogue
# (deep_zoom_points.py:286-318), truncated to 128 digits: zoom 1e100 views
# a 1e-100 neighbourhood, far above the truncation.
TRICORN_POINTS = {
    "branching-junction": ("-1.09692656551687871996070994567030699382726636856135252864063809062918671538233200238925924425047165728705144577721323870097123830",
        "0.15924841985780843944738929716460993014535067284697776610028746670435199873283411089165347426523171986788898191180737561810646406"),
    "preperiodic-fork": ("-1.22224542629255890298019006722054650655408214278112662131889956070689001518195800647023583458966697095257488470249030425971271679",
        "0.18411010266019595907154534797192084995346033751489213057770255983684726093175172244070707373577615831859868664198973001227361970"),
    "antler-junction": ("-0.89496367244425123674835666927614371605103297638091872536445395141229693274347387622099879227927795268786190095948036005167118619",
        "0.16233379791644897545645068663668796188752900772412238190448090899609106894212265824125641030767828484157713045472396312957550898"),
}


def tricorn_field(point: str = "branching-junction", width: int = 3840, height: int = 2160, zoom: str = "1e100",
                  max_iter: int = 200000):
    """A tricorn field centred on a catalogue boundary root, 480-bit orbit,
    default render options."""
    x, y = TRICORN_POINTS[point]
    return dict(name=f"tricorn_{point}_{width}x{height}_zoom{zoom}_cap{max_iter}", formula=3, julia=("0", "0"),
                x=x, y=y, zoom=zoom, width=width, height=height, max_iter=max_iter, bits=480)


def atlas_tile(width, height, max_iter, seed, centered=False):
    """A synthetic scalar atlas tile: smooth escape ramps with an interior
    block, an interior left edge and one glitch sentinel; with `centered`,
    stored relative to the tile's own cap (value - max_iter) as the
    production atlas encoding does."""
    import numpy as np
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    field = (300.0 + 40.0 * np.sin(x / 6.0) * np.cos(y / 5.0) + 1.3 * x - 0.7 * y
             + rng.normal(0, 1.5, (height, width)))
    field[height // 3:height // 3 + 7, width // 2:width // 2 + 9] = max_iter   # interior block
    field[:, :3] = max_iter                                                  # interior left edge
    field[height - 2, width - 5] = np.nan                                    # glitch sentinel
    field = field.astype(np.float64)
    if centered:
        field = field - max_iter
    return field.astype(np.float32)
