"""Writes tests/golden/<repair case>.npz: the glitch-repair driver
(fav_repair.repair_glitches) run over the reference's own renderer.cpp.

Run in a container where /root/reference exists, after ``make -C oracle
ref``.  The strict initial field and every point render of the repair come
from renderer.cpp's scalar backend (backend 0).  The generator checks that
none of those point renders counted a glitch, so the scalar backend's exact
MPFR re-run of glitched points (renderer.cpp:12944-12966), which the GPU
point path does not make, never decided a value.

    python tests/golden/make_golden_repair.py
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent.parent / "fractal-audio-viz_amd"))

import fav_repair  # noqa: E402
import native  # noqa: E402
import repair_io  # noqa: E402
from repair_cases import CASES  # noqa: E402


def initial_field(lib, case, backend):
    """The strict field the repair starts from."""
    z = fav_repair.zoom_text(case["log10_zoom"])
    bits = fav_repair.native_precision_bits(case["x"], case["y"], case["log10_zoom"])
    opts = native.RenderOptions.defaults(backend=backend, **case["initial"])
    ref = native.AbiReference(lib, case["x"], case["y"], z, case["max_iter"], bits, repair_io.IMAGE_SERIES_ORDER,
                              opts)
    field, _, _ = ref.render(case["width"], case["height"], z, case["max_iter"], case["series_order"],
                             case["series_block"], opts)
    return field


def repair(case, field, factory, backend):
    diagnostics = {}
    repaired = fav_repair.repair_glitches(field, log10_zoom=case["log10_zoom"], x_center=case["x"],
                                          y_center=case["y"], max_iter=case["max_iter"],
                                          series_order=case["series_order"], series_block=case["series_block"],
                                          reference_factory=factory, backend=backend, diagnostics=diagnostics)
    return repaired, diagnostics


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    for case in CASES:
        started = time.time()
        field = initial_field(lib, case, 0)
        repair_io.AbiSecondary.glitches = 0
        repaired, diagnostics = repair(case, field, lambda *a: repair_io.AbiSecondary(lib, *a, backend=0), 0)
        if repair_io.AbiSecondary.glitches:
            raise SystemExit(f"{case['name']}: {repair_io.AbiSecondary.glitches} glitches in the repair's point "
                             "renders would take the scalar backend's exact re-run")
        meta = dict(case, diagnostics=diagnostics)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(meta)), initial=field,
                            repaired=repaired)
        print(case["name"], field.shape, "nan", int(np.isnan(field).sum()), diagnostics,
              f"{time.time() - started:.1f}s", flush=True)


if __name__ == "__main__":
    main()
