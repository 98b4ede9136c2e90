"""Writes tests/golden/kfp_planes_*.npz: the reference's own plane-aware KFP
colour pass (fractal_colourise_kfp_planes of oracle/_ref/mandelbrot.so,
built from renderer.cpp) on planes derived from the field golden vectors.
The inputs are regenerated by tests/kfp_planes_io.py; the fixtures keep
the RGB output.

    python tests/golden/make_golden_kfp_planes.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import kfp_planes_io  # noqa: E402
import native  # noqa: E402
from kfp_planes_cases import ATLAS_CASES, CASES, CROP_CASES, RECT_CASES  # noqa: E402


def main():
    ref = native.load_reference()
    if ref is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    native.bind_kfp(ref)
    oracle = native.load_oracle()
    for case in CASES:
        field, planes, lut, opts, max_iter = kfp_planes_io.inputs(case, oracle)
        rgb = kfp_planes_io.colourise(ref, field, planes, lut, opts, max_iter)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(case)), rgb=rgb)
        print(case["name"], field.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)
    for case in RECT_CASES:
        parent, child, lut, opts = kfp_planes_io.rect_inputs(case, oracle)
        rgb = kfp_planes_io.rect_colourise(ref, case, parent, child, lut, opts)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(case)), rgb=rgb)
        print(case["name"], rgb.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)
    for case in CROP_CASES:
        field, lut, opts = kfp_planes_io.crop_inputs(case, oracle)
        rgb = kfp_planes_io.crop_colourise(ref, case, field, lut, opts)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(case)), rgb=rgb)
        print(case["name"], rgb.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)
    for case in ATLAS_CASES:
        parent, child, lut, opts = kfp_planes_io.atlas_inputs(case, oracle)
        rgb = kfp_planes_io.atlas_colourise(ref, case, parent, child, lut, opts)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(case)), rgb=rgb)
        print(case["name"], rgb.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)


if __name__ == "__main__":
    main()
