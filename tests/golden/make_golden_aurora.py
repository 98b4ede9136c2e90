"""Writes tests/golden/aurora_*.npz from the reference's own renderer.cpp.

Run in a container where /root/reference exists, after ``make -C oracle
ref``.  Each fixture holds a checksum of the case's regenerated inputs and
the function's output.

    python tests/golden/make_golden_aurora.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import aurora_io  # noqa: E402
import native  # noqa: E402
from aurora_cases import CASES  # noqa: E402


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    aurora_io.bind(lib)
    for case in CASES:
        status, out = aurora_io.run(lib, case)
        if status != 0:
            raise SystemExit(f"{case['name']}: {lib.fractal_last_error().decode()}")
        meta = dict(case, inputs_sha256=aurora_io.checksum(aurora_io.inputs(case)))
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(meta)), output=out)
        print(case["name"], out.shape, out.dtype, flush=True)


if __name__ == "__main__":
    main()
