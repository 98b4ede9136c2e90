"""Writes tests/golden/kfp_<case>.npz: the reference's own KFP colour transfer
(oracle/_ref/mandelbrot.so, built from renderer.cpp) applied to golden or
seeded scalar fields.  The palette LUT is built by the oracle's restatement
of visualizer.py:13492-13521 (pinned separately by its closed form).

    python tests/golden/make_golden_kfp.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import native  # noqa: E402
from kfp_cases import CASES, KALLES_DEFAULT_STOPS, seeded_stops, synthetic_field  # noqa: E402


def case_inputs(oracle, case):
    name, source, max_iter, stops_kind, options = case
    if source == "synthetic":
        field = synthetic_field(max_iter=max_iter)
    else:
        with np.load(HERE / f"{source}.npz", allow_pickle=False) as data:
            field = data["field"].astype(np.float32)
    stops = np.asarray(KALLES_DEFAULT_STOPS if stops_kind == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    opts = native.KfpOptions.make(**options)
    return field, lut, opts


def main():
    ref = native.load_reference()
    if ref is None:
        raise SystemExit("oracle/_ref/mandelbrot.so is missing: run make -C oracle ref")
    native.bind_kfp(ref)
    oracle = native.load_oracle()
    for case in CASES:
        field, lut, opts = case_inputs(oracle, case)
        rgb = native.colourise_abi(ref, field, case[2], opts, lut)
        np.savez_compressed(HERE / f"kfp_{case[0]}.npz", field=field, lut=lut, rgb=rgb,
                            max_iter=np.array(case[2]), options=np.array(json.dumps(opts.as_dict())))
        print(case[0], field.shape, "mean rgb", rgb.reshape(-1, 3).mean(axis=0).round(1), flush=True)


if __name__ == "__main__":
    main()
