"""Writes tests/golden/atlas_<case>.npz: the reference's own raw KFP atlas
colour pass (oracle/_ref/mandelbrot.so, built from renderer.cpp) on the
seeded parent/child fields of atlas_cases.py.

    python tests/golden/make_golden_atlas.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parents[1] / "fractal-audio-viz_amd"))

import atlas_io  # noqa: E402
import native  # noqa: E402
from atlas_cases import CASES  # noqa: E402


def main():
    ref = native.load_reference()
    if ref is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    fn = atlas_io.bind_atlas(ref)
    oracle = native.load_oracle()
    for case in CASES:
        parent, child, lut, opts = atlas_io.case_inputs(oracle, case)
        status, rgb = atlas_io.atlas_call(fn, case, parent, child, lut, opts)
        if status != 0:
            raise SystemExit(f"{case['name']}: {ref.fractal_last_error().decode()}")
        np.savez_compressed(HERE / f"{case['name']}.npz", rgb=rgb, meta=np.array(json.dumps(case)))
        print(case["name"], rgb.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)


if __name__ == "__main__":
    main()
