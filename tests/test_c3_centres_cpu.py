"""The seeded near-boundary centre generator of the 2560x1440 precision
fields (fav_workloads.nucleus_centre) against closed forms, independent
iteration and its committed output (tests/golden/c3_centres.json)."""
import json
from pathlib import Path

import mpmath as mp
import numpy as np
import pytest

import fav_workloads

CENTRES = json.loads((Path(__file__).resolve().parent / "golden" / "c3_centres.json").read_text())


def iterate(formula, a, b, n):
    """z_1 .. z_n of the map at c = (a, b), written out independently of the
    generator: Mandelbrot z^2 + c, burning ship (|x| + i|y|)^2 + c."""
    x = y = mp.mpf(0)
    out = []
    for _ in range(n):
        if formula == 2:
            x, y = abs(x), abs(y)
        x, y = x * x - y * y + a, 2 * x * y + b
        out.append((x, y))
    return out


@pytest.mark.parametrize("key", sorted(CENTRES))
def test_generator_reproduces_committed_centres(key):
    entry = CENTRES[key]
    again = fav_workloads.nucleus_centre(entry["formula"], entry["seed"], entry["zoom_exponent"])
    for field in ("x", "y", "nucleus_x", "nucleus_y", "period", "size"):
        assert again[field] == entry[field], field


@pytest.mark.parametrize("key", sorted(CENTRES))
def test_committed_centres_are_nuclei_of_their_period(key):
    """|z_p| at the 120-digit nucleus is below 1e-10 of its value at the
    centre 10^-(zoom + 2) away (the root is far closer than the offset), no
    earlier z_k vanishes (exact period p), p lies in [50, 500], the size
    estimate is 1/480 to 1/8 of the view height, and the centre sits
    10^-(zoom + 2) from the nucleus."""
    e = CENTRES[key]
    zoom = e["zoom_exponent"]
    with mp.workdps(150):
        a, b = mp.mpf(e["nucleus_x"]), mp.mpf(e["nucleus_y"])
        orbit = iterate(e["formula"], a, b, e["period"])
        at_centre = iterate(e["formula"], mp.mpf(e["x"]), mp.mpf(e["y"]), e["period"])[-1]
        assert 50 <= e["period"] <= 500
        assert mp.sqrt(orbit[-1][0] ** 2 + orbit[-1][1] ** 2) < 1e-10 * mp.sqrt(at_centre[0] ** 2 + at_centre[1] ** 2)
        assert min(mp.sqrt(x * x + y * y) for x, y in orbit[:-1]) > mp.mpf(10) ** -20
        view = mp.mpf("2.8") * mp.mpf(10) ** -zoom
        assert view / 480 <= mp.mpf(e["size"]) * (1 + 1e-5) and mp.mpf(e["size"]) <= view / 8
        offset = mp.sqrt((mp.mpf(e["x"]) - a) ** 2 + (mp.mpf(e["y"]) - b) ** 2)
        assert abs(offset / mp.mpf(10) ** -(zoom + 2) - 1) < 1e-15


def test_mandelbrot_newton_finds_the_period_2_and_3_nuclei():
    """Period 2: c = -1.  Period 3: the roots of c^3 + 2c^2 + c + 1
    (z_3 = c^4 + 2c^3 + c^2 + c = c (c^3 + 2c^2 + c + 1))."""
    with mp.workdps(60):
        tol = mp.mpf(10) ** -50
        a, b = fav_workloads._newton(0, mp.mpf("-0.9"), mp.mpf("0.05"), 2, tol)
        assert abs(a + 1) < 1e-50 and abs(b) < 1e-50
        for root in np.roots([1, 2, 1, 1]):
            start = (mp.mpf(root.real) + mp.mpf("0.01"), mp.mpf(root.imag) - mp.mpf("0.01"))
            a, b = fav_workloads._newton(0, *start, 3, tol)
            assert abs(mp.mpc(a, b) - mp.mpc(root.real, root.imag)) < 1e-12


def test_size_estimate_closed_forms():
    """Main cardioid (period 1, c = 0): 1.  Period-2 disk (c = -1): the
    estimate 1 / |beta lambda^2| with lambda = 2 z_1 = -2, beta = 1 + 1/lambda
    = 1/2, i.e. 1/2 (the disk's diameter)."""
    with mp.workdps(40):
        assert fav_workloads.nucleus_properties(0, mp.mpf(0), mp.mpf(0), 1)[2] == 1
        assert abs(fav_workloads.nucleus_properties(0, mp.mpf(-1), mp.mpf(0), 2)[2] - mp.mpf("0.5")) < 1e-30


@pytest.mark.parametrize("start,period", [(("0.3", "-0.5"), 6), (("-1.0", "-0.9"), 4)])
def test_burning_ship_newton_carries_the_fold_signs(start, period):
    """Off-axis burning-ship nuclei whose orbits cross the axes (x y < 0
    on several steps), so the real Jacobian needs each fold's sign: the
    root found is a nucleus of exact period `period` by independent
    iteration."""
    with mp.workdps(60):
        root = fav_workloads._newton(2, mp.mpf(start[0]), mp.mpf(start[1]), period, mp.mpf(10) ** -50)
        assert root is not None
        orbit = iterate(2, root[0], root[1], period)
        assert sum(1 for x, y in orbit[:-1] if x * y < 0) >= 1
        assert mp.sqrt(orbit[-1][0] ** 2 + orbit[-1][1] ** 2) < 1e-45
        assert min(mp.sqrt(x * x + y * y) for x, y in orbit[:-1]) > 1e-3
        assert abs(root[1]) > 0.05


def test_anchor_refinement_reaches_the_misiurewicz_points():
    """c = i satisfies z_4 = z_2 exactly; the real anchor is the root of
    c^3 + 2c^2 + 2c + 2 (z_3 = -z_2)."""
    with mp.workdps(80):
        a, b = fav_workloads._anchor("1e-10", "1.0000000001", 2, 2)
        assert abs(a) < mp.mpf(10) ** -70 and abs(b - 1) < mp.mpf(10) ** -70
        a, _ = fav_workloads._anchor("-1.5436890126920764", "0", 3, 1)
        assert abs(a ** 3 + 2 * a ** 2 + 2 * a + 2) < mp.mpf(10) ** -70
