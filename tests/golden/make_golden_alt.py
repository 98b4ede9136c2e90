"""Writes tests/golden/alt_*.npz from the reference's own renderer.cpp.

Run where /root/reference exists, after ``make -C oracle ref``:

    python tests/golden/make_golden_alt.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import golden_io  # noqa: E402
import native  # noqa: E402
from alt_cases import CASES  # noqa: E402


def main():
    lib = native.load_reference()
    if lib is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    only = set(sys.argv[1:])
    for case in (c for c in CASES if not only or c["name"] in only):
        field, stats = golden_io.render_alt_abi(lib, case, backend=0)
        np.savez_compressed(HERE / f"{case['name']}.npz", meta=np.array(json.dumps(case)),
                            field=field.astype(np.float32), stats=stats.astype(np.uint64))
        fin = field[np.isfinite(field)]
        print(case["name"], "nan", int(np.isnan(field).sum()), "interior", int((field >= case["max_iter"]).sum()),
              "range %.2f %.2f" % (fin.min(), fin.max()), flush=True)


if __name__ == "__main__":
    main()
