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


# ---- near-boundary centres: Newton refinement to a minibrot nucleus ------
#
# A centre is found in three seeded steps, at `digits` + 30 decimal digits
# of working precision (mpmath):
#   1. a start point c0 at a seeded distance 10^-k and angle from a seeded
#      anchor, a boundary point where the map is expanding (a Misiurewicz
#      point), so that the minibrots of period p near it shrink
#      geometrically with p;
#   2. the "partials" of c0: the iterations n <= 500 at which |z_n| reaches
#      a new minimum.  Each is the period of a nucleus whose atom domain
#      holds c0, so Newton's method started at c0 converges to it (the atom
#      domain method of C. Heiland-Allen);
#   3. for the partials p in [50, 500] in increasing order, Newton's method
#      on z_p(c) = 0 from c0, kept when it converges to a nucleus of exact
#      period p whose estimated minibrot size lies between 1/480 and 1/8 of
#      the view height (2.8 / zoom): the minibrot shows in the view,
#      surrounded by its decorations.
# The centre is that nucleus plus a seeded 10^-(zoom exponent + 2) offset.
# The anchors are chosen for slow shrinking (about 1.3x to 2.4x per period
# step), so that both zoom 1e30 and zoom 1e100 have sizes in range at
# periods in [50, 500]; c = -2, whose orbit lands on the escape circle
# |z| = 2, is left out.
# The burning ship (x' = x^2 - y^2 + a, y' = 2|x||y| + b, renderer.cpp:
# 268-279 and 10837-10846) is not holomorphic, so its Newton step solves
# the real 2x2 system of d(x_p, y_p)/d(a, b), whose chain rule carries the
# sign s = sign(x) sign(y) of each fold.  On the real axis it is the real
# Mandelbrot map, so its real Misiurewicz points serve as anchors too.

# (x, y, preperiod k, period q): z_{k+q} = z_k, refined by _anchor to the
# working precision (a start point's offset from an anchor that is exact
# only to 17 digits would be swamped by the anchor's own error).  The
# principal Misiurewicz point of the 1/3 limb and its mirror image, the
# real point M_{3,1} (root of c^3 + 2c^2 + 2c + 2) and c = i.
MANDELBROT_ANCHORS = (("-0.10109636384562216", "0.9562865108091415", 4, 1),
                      ("-0.10109636384562216", "-0.9562865108091415", 4, 1),
                      ("-1.5436890126920764", "0", 3, 1), ("0", "1", 2, 2))
SHIP_ANCHORS = (("-1.5436890126920764", "0", 3, 1),)


def _anchor(x, y, preperiod, period):
    """The Misiurewicz point z_{k+q}(c) = z_k(c) nearest (x, y), by Newton's
    method on the Mandelbrot map at the current precision."""
    import mpmath as mp
    c = mp.mpc(mp.mpf(x), mp.mpf(y))
    for _ in range(200):
        z, dz, zk, dzk = mp.mpc(0), mp.mpc(0), None, None
        for n in range(1, preperiod + period + 1):
            z, dz = z * z + c, 2 * z * dz + 1
            if n == preperiod:
                zk, dzk = z, dz
        step = (z - zk) / (dz - dzk)
        c -= step
        if abs(step) < mp.mpf(10) ** -(mp.mp.dps - 5):
            break
    return c.real, c.imag


def _ship_step(x, y, a, b):
    return x * x - y * y + a, 2 * abs(x) * abs(y) + b


def _partials(formula, a, b, limit=500):
    """[(n, size estimate)] for the partials n of c0 = (a, b): the
    nucleus_properties size formula evaluated along c0's own orbit, which
    shadows the nucleus' orbit up to n, so that Newton's method runs only
    on the partials whose minibrot can be in range."""
    import mpmath as mp
    x = y = xa = xb = ya = yb = mp.mpf(0)
    best, gain, out = mp.inf, mp.mpf(1), []
    for n in range(1, limit + 1):
        s = (-1 if (x < 0) != (y < 0) else 1) if formula == 2 else 1
        xa, xb, ya, yb = (2 * x * xa - 2 * y * ya + 1, 2 * x * xb - 2 * y * yb,
                          2 * s * (y * xa + x * ya), 2 * s * (y * xb + x * yb) + 1)
        x, y = _ship_step(x, y, a, b) if formula == 2 else (x * x - y * y + a, 2 * x * y + b)
        r = x * x + y * y
        if r < best:
            best = r
            dzdc = mp.sqrt(abs(xa * yb - xb * ya))
            out.append((n, 1 / (dzdc * gain) if dzdc > 0 else mp.inf))
        if r > 4:
            break
        gain *= 2 * mp.sqrt(r)
    return out


def _newton(formula, a, b, period, tolerance, steps=80):
    """Newton's method on z_period(c) = 0; (a, b) of the converged root or
    None.  The Jacobian d(x, y)/d(a, b) is carried as four reals (for the
    Mandelbrot map it is the Cauchy-Riemann matrix of dz/dc)."""
    import mpmath as mp
    for _ in range(steps):
        x = y = xa = xb = ya = yb = mp.mpf(0)
        for _ in range(period):
            if formula == 2:
                s = -1 if (x < 0) != (y < 0) else 1
                xa, xb, ya, yb = (2 * x * xa - 2 * y * ya + 1, 2 * x * xb - 2 * y * yb,
                                  2 * s * (y * xa + x * ya), 2 * s * (y * xb + x * yb) + 1)
                x, y = _ship_step(x, y, a, b)
            else:
                xa, xb, ya, yb = (2 * x * xa - 2 * y * ya + 1, 2 * x * xb - 2 * y * yb,
                                  2 * (y * xa + x * ya), 2 * (y * xb + x * yb) + 1)
                x, y = x * x - y * y + a, 2 * x * y + b
            if abs(x) + abs(y) > 1e50:
                return None
        det = xa * yb - xb * ya
        if det == 0:
            return None
        da, db = (yb * x - xb * y) / det, (xa * y - ya * x) / det
        a, b = a - da, b - db
        if max(abs(da), abs(db)) < tolerance:
            return a, b
    return None


def nucleus_properties(formula, a, b, period):
    """(min |z_n| over 0 < n < period, |z_period|, size estimate) of a
    candidate nucleus.  The size is 1 / (|dz_p/dc| prod_{n<p} 2|z_n|): for
    the Mandelbrot map this is the minibrot size estimate 1 / |beta lambda^2|
    (lambda = prod 2 z_n, dz_p/dc = beta lambda); for the burning ship
    |dz_p/dc| is the square root of the Jacobian determinant."""
    import mpmath as mp
    x = y = xa = xb = ya = yb = mp.mpf(0)
    smallest, gain = mp.inf, mp.mpf(1)
    for n in range(1, period + 1):
        s = (-1 if (x < 0) != (y < 0) else 1) if formula == 2 else 1
        xa, xb, ya, yb = (2 * x * xa - 2 * y * ya + 1, 2 * x * xb - 2 * y * yb,
                          2 * s * (y * xa + x * ya), 2 * s * (y * xb + x * yb) + 1)
        x, y = _ship_step(x, y, a, b) if formula == 2 else (x * x - y * y + a, 2 * x * y + b)
        if n < period:
            modulus = mp.sqrt(x * x + y * y)
            smallest = min(smallest, modulus)
            gain *= 2 * modulus
    scale = mp.sqrt(abs(xa * yb - xb * ya)) * gain
    return smallest, mp.sqrt(x * x + y * y), 1 / scale if scale != 0 else mp.inf


def nucleus_centre(formula: int, seed: int, zoom_exponent: int, digits: int = 120, attempts: int = 200):
    """A seeded near-boundary centre for a view at zoom 10^zoom_exponent:
    the Newton-refined nucleus of a period-p minibrot (p in [50, 500],
    estimated size 1/480 to 1/8 of the view height) plus a seeded
    10^-(zoom_exponent+2) offset, as `digits`-digit decimal text.  formula
    0 is the Mandelbrot map, 2 the burning ship.  Raises if no start point
    within `attempts` converges."""
    import mpmath as mp
    if formula not in (0, 2):
        raise ValueError("nucleus centres are defined for the Mandelbrot map (0) and the burning ship (2)")
    rng = random.Random(f"nucleus-{formula}-{seed}-{zoom_exponent}")
    anchors = SHIP_ANCHORS if formula == 2 else MANDELBROT_ANCHORS
    with mp.workdps(digits + 30):
        tolerance = mp.mpf(10) ** -(digits + 15)
        view = mp.mpf("2.8") * mp.mpf(10) ** -zoom_exponent
        smallest_size, largest_size = view / 480, view / 8
        exact = {(a[0], a[1]): _anchor(*a) for a in anchors}
        for attempt in range(attempts):
            ax, ay, preperiod, period = anchors[rng.randrange(len(anchors))]
            a_exact, b_exact = exact[(ax, ay)]
            depth = rng.uniform(0.6 * zoom_exponent, 1.1 * zoom_exponent)
            angle = 2 * mp.pi * rng.random()
            a0 = a_exact + mp.mpf(10) ** -depth * mp.cos(angle)
            b0 = b_exact + mp.mpf(10) ** -depth * mp.sin(angle)
            for period, estimate in _partials(formula, a0, b0):
                if period < 50 or estimate > 100 * largest_size:
                    continue
                if estimate < smallest_size / 100:
                    break
                root = _newton(formula, a0, b0, period, tolerance)
                if root is None:
                    continue
                smallest, residual, size = nucleus_properties(formula, root[0], root[1], period)
                if smallest < mp.mpf(10) ** -40:
                    continue   # the nucleus of a smaller period
                if size > largest_size:
                    continue   # too large: try the next partial
                if size < smallest_size:
                    break      # later partials are smaller still: next start point
                offset_angle = 2 * mp.pi * rng.random()
                step = mp.mpf(10) ** -(zoom_exponent + 2)
                x = root[0] + step * mp.cos(offset_angle)
                y = root[1] + step * mp.sin(offset_angle)
                return dict(x=mp.nstr(x, digits, strip_zeros=False), y=mp.nstr(y, digits, strip_zeros=False),
                            nucleus_x=mp.nstr(root[0], digits, strip_zeros=False),
                            nucleus_y=mp.nstr(root[1], digits, strip_zeros=False), period=period,
                            size=mp.nstr(size, 6), attempt=attempt, anchor=[ax, ay])
    raise RuntimeError(f"no nucleus found for formula {formula}, seed {seed}, zoom 1e{zoom_exponent}")


def c3_bits(zoom_exponent: int) -> int:
    """Orbit precision of a near-boundary field: 64 bits above the zoom's,
    rounded up to a multiple of 32."""
    import math
    return 32 * math.ceil((64 + zoom_exponent * math.log2(10)) / 32)


def near_boundary_field(formula: int, seed: int, zoom_exponent: int, width: int = 2560, height: int = 1440,
                        max_iter: int = 100000, centre=None):
    """A 2560x1440, cap-100000 field at zoom 10^zoom_exponent around a
    nucleus_centre: orbit precision c3_bits, glitch rebase on
    (allow_recovery), the reference's default image series, BLA order 3 with
    block 256 for the Mandelbrot map (the alternate formulas take no BLA
    order).  The orbit near a nucleus returns close to 0 every period, which
    drives the reference's glitch handling: with strict rendering and no
    rebase most of such a view is NaN."""
    c = centre if centre is not None else nucleus_centre(formula, seed, zoom_exponent)
    name = "mandelbrot" if formula == 0 else "burning_ship"
    return dict(name=f"{name}_{width}x{height}_zoom1e{zoom_exponent}_cap{max_iter}_seed{seed}", formula=formula,
                julia=("0", "0"), x=c["x"], y=c["y"], zoom=f"1e{zoom_exponent}", width=width, height=height,
                max_iter=max_iter, bits=c3_bits(zoom_exponent), image_series_order=8,
                series_order=3 if formula == 0 else 8, series_block=256 if formula == 0 else 32,
                options=dict(allow_recovery=1), period=c["period"])


def misiurewicz_field(width: int = 1920, height: int = 1080, zoom_exponent: int = 100, max_iter: int = 100000,
                      seed: int = 1):
    """A boundary-rich Mandelbrot field whose reference orbit never nears 0:
    the principal Misiurewicz point of the 1/3 limb (refined to 150 digits)
    plus a seeded 10^-(zoom_exponent+1) offset.  The view shows the spiral
    around that preperiodic point; every pixel escapes (about 806..858
    iterations at zoom 1e100), none glitches, and the counts equal
    arbitrary-precision iteration.  Orbit precision c3_bits; options as the
    headline field's planes and Kalles colouring, R = 2."""
    import mpmath as mp
    rng = random.Random(f"misiurewicz-{seed}-{zoom_exponent}")
    with mp.workdps(150):
        a, b = _anchor(*MANDELBROT_ANCHORS[0])
        angle = 2 * mp.pi * rng.random()
        step = mp.mpf(10) ** -(zoom_exponent + 1)
        x = mp.nstr(a + step * mp.cos(angle), 120, strip_zeros=False)
        y = mp.nstr(b + step * mp.sin(angle), 120, strip_zeros=False)
    return dict(name=f"misiurewicz_{width}x{height}_zoom1e{zoom_exponent}_cap{max_iter}", x=x, y=y,
                zoom=f"1e{zoom_exponent}", width=width, height=height, max_iter=max_iter,
                bits=c3_bits(zoom_exponent), image_series_order=8, series_order=3, series_block=256,
                options=dict(allow_recovery=1), planes=True)
