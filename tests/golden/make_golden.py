"""Writes tests/golden/<case>.npz from the reference's own renderer.cpp.

Run in a container where /root/reference exists, after
``make -C oracle ref`` built oracle/_ref/mandelbrot.so from the reference
sources (with -ffp-contract=off).  The fixtures hold inputs and outputs only:
the float32 field, the six render planes, and the RenderStats counters.

    python tests/golden/make_golden.py [case names, default all]
"""
import ctypes
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import native  # noqa: E402
from cases import CASES  # noqa: E402


def render_case(lib, case):
    opts = native.RenderOptions.defaults(**case["options"])
    root = native.AbiReference(lib, case["x"], case["y"], case["ref_zoom"], case["max_iter"],
                               case["bits"], case["image_series_order"], opts)
    target = root
    if case["clone_zoom"]:
        # clones need a reusable root (renderer.cpp:13707-13710)
        handle = lib.fractal_create_reference_reusable_options(
            case["x"].encode(), case["y"].encode(), case["ref_zoom"].encode(), case["max_iter"],
            case["bits"], case["image_series_order"], ctypes.byref(opts))
        assert handle, lib.fractal_last_error()
        clone = lib.fractal_clone_reference(ctypes.c_void_p(handle), case["clone_zoom"].encode())
        assert clone, lib.fractal_last_error()
        lib.fractal_destroy_reference(ctypes.c_void_p(handle))
        target = native.AbiReference.__new__(native.AbiReference)
        target.lib, target.handle = lib, clone
    field, planes, stats = target.render(case["width"], case["height"], case["zoom"],
                                         case["max_iter"], case["series_order"],
                                         case["series_block"], opts, planes=True, stats=True)
    return field, planes, stats


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    only = set(sys.argv[1:])
    for case in (c for c in CASES if not only or c["name"] in only):
        field, planes, stats = render_case(lib, case)
        meta = {k: v for k, v in case.items()}
        np.savez_compressed(
            HERE / f"{case['name']}.npz",
            meta=np.array(json.dumps(meta)),
            field=field.astype(np.float32),
            orbit_iteration=planes.orbit_iteration.astype(np.int64),
            phase=planes.phase, de_x=planes.de_x, de_y=planes.de_y,
            test1=planes.test1, test2=planes.test2,
            stats=stats.astype(np.uint64))
        print(case["name"], field.shape, "nan", int(np.isnan(field).sum()),
              "bla", int(stats[2]), "exact", int(stats[5]), flush=True)


if __name__ == "__main__":
    main()
