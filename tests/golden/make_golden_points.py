"""Writes tests/golden/<point case>.npz from the reference's own renderer.cpp.

Run in a container where /root/reference exists, after
``make -C oracle ref`` built oracle/_ref/mandelbrot.so from the reference
sources (with -ffp-contract=off).  Each fixture holds the case's point
inputs (mantissa pairs and exponents), the float32 output of
fractal_render_points on the scalar backend (backend 0) and the
RenderStats counters.

    python tests/golden/make_golden_points.py
"""
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import native  # noqa: E402
from point_cases import CASES, make_points  # noqa: E402

GLITCH_COUNT = 18  # RenderStats.glitch_count
EXACT_REPAIR_LIMIT = 8192  # MAX_EXACT_POINT_REPAIR_PIXELS, renderer.cpp:76


def render_points(lib, case, backend=0, threads=0):
    """One point render through a renderer.h-speaking library."""
    re, im, e = make_points(case)
    opts = native.RenderOptions.defaults(backend=backend, **case["options"])
    ref = native.AbiReference(lib, case["x"], case["y"], case["ref_zoom"], case["max_iter"], case["bits"],
                              case["image_series_order"], opts, formula=case["formula"], julia=case["julia"])
    out = np.empty(len(re), dtype=np.float32)
    lib.fractal_set_stats_enabled(1)
    status = lib.fractal_render_points(
        out.ctypes.data_as(native.c_float_p), len(re), case["zoom"].encode(),
        re.ctypes.data_as(native.c_double_p), im.ctypes.data_as(native.c_double_p),
        e.ctypes.data_as(native.c_int32_p), ctypes.c_void_p(ref.handle), case["max_iter"], threads,
        case["series_order"], case["series_block"], ctypes.byref(opts))
    stats = np.zeros(39, dtype=np.uint64)
    lib.fractal_get_last_stats_ex(stats.ctypes.data_as(native.c_uint64_p), 39)
    lib.fractal_set_stats_enabled(0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return (re, im, e), out, stats


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    for case in CASES:
        started = time.time()
        (re, im, e), out, stats = render_points(lib, case)
        strict = case["options"].get("strict", 1) != 0
        # only the Mandelbrot loop repairs (alternates report no glitch counter)
        if case["formula"] == 0 and strict and len(re) <= EXACT_REPAIR_LIMIT and stats[GLITCH_COUNT] != 0:
            raise SystemExit(f"{case['name']}: {int(stats[GLITCH_COUNT])} glitches would take the "
                             "scalar backend's exact repair")
        meta = {k: v for k, v in case.items() if k != "fixed"}
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(meta)),
                            real=re, imag=im, exponent=e, field=out, stats=stats)
        print(case["name"], len(out), "nan", int(np.isnan(out).sum()), "glitch", int(stats[GLITCH_COUNT]),
              "bla", int(stats[2]), "exact", int(stats[5]), f"{time.time() - started:.1f}s", flush=True)


if __name__ == "__main__":
    main()
