"""Writes tests/golden/direct_*.npz from the reference's own renderer.cpp
(scalar backend, use_perturbation = 0).

Run where /root/reference exists, after ``make -C oracle ref``:

    python tests/golden/make_golden_direct.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import native  # noqa: E402
from direct_cases import CASES  # noqa: E402


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    for case in CASES:
        field, planes = native.render_direct_abi(lib, case, backend=0)
        arrays = {"field": field, "meta": np.array(json.dumps(case))}
        if planes is not None:
            arrays.update(planes.as_dict())
        np.savez_compressed(HERE / f"{case['name']}.npz", **arrays)
        print(case["name"], "nan", int(np.isnan(field).sum()), "interior",
              int((field == case["max_iter"] - case["options"].get("output_bias", 0.0)).sum()))


if __name__ == "__main__":
    main()
