"""Writes tests/golden/repair_visualizer_<case>.npz: the reference's own
glitch-repair function (visualizer.py:8215-9252,
`_atlas_glitch_reference_field`) run on the repair cases of repair_cases.py,
starting from the same strict initial field as the case's fixture, over the
reference's renderer.cpp built from source (oracle/_ref/mandelbrot.so,
scalar backend 0).  These fixtures pin fav_repair's driver against the
reference's: tests/test_repair_cpu.py compares them with the driver's
output.

Run only in a container where /root/reference exists, after
``make -C oracle ref``:

    python tests/golden/make_golden_repair_visualizer.py
"""
import json
import os
import sys
from pathlib import Path

import numpy as np  # before the reference library: the system libstdc++ first

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
os.environ["MANDELBROT_LIBRARY"] = str(ROOT / "oracle" / "_ref" / "mandelbrot.so")
sys.path.insert(0, "/root/reference")
sys.path.insert(0, str(HERE))

import visualizer  # noqa: E402
from repair_cases import CASES  # noqa: E402


def main():
    library = visualizer._get_native_library()
    if library is None:
        raise SystemExit("oracle/_ref/mandelbrot.so did not load: run make -C oracle ref")
    for case in CASES:
        fixture = np.load(HERE / f"{case['name']}.npz", allow_pickle=False)
        reference, _ = visualizer._create_native_reference(case["x"], case["y"], case["max_iter"], case["log10_zoom"],
                                                           case["series_order"])
        diagnostics = {}
        field = visualizer._atlas_glitch_reference_field(
            render_width=case["width"], render_height=case["height"], log10_zoom=case["log10_zoom"],
            x_center=case["x"], y_center=case["y"], max_iter=case["max_iter"], series_order=case["series_order"],
            series_block=case["series_block"], native_threads=8, native_library=library, native_backend=0,
            native_reference=reference, fallback_field=None, fallback_zoom_factor=2.0, fallback_max_iter=None,
            allow_recovery=False, diagnostics=diagnostics, initial_field=fixture["initial"])
        np.savez_compressed(HERE / f"repair_visualizer_{case['name'][len('repair_'):]}.npz",
                            meta=np.array(json.dumps(dict(name=case["name"], diagnostics=diagnostics))),
                            field=np.asarray(field, dtype=np.float32))
        print(case["name"], diagnostics, flush=True)


if __name__ == "__main__":
    main()
