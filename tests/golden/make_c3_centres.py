"""Writes tests/golden/c3_centres.json: the seeded near-boundary centres of
the 2560x1440 precision fields (Mandelbrot and burning ship at zooms 1e30
and 1e100), from fav_workloads.nucleus_centre with seed 1.

    python tests/golden/make_c3_centres.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "fractal-audio-viz_amd"))

import fav_workloads  # noqa: E402

SEED = 1
FIELDS = [(0, 30), (0, 100), (2, 30), (2, 100)]


def main():
    out = {}
    for formula, zoom_exponent in FIELDS:
        centre = fav_workloads.nucleus_centre(formula, SEED, zoom_exponent)
        out[f"{formula}_{zoom_exponent}"] = dict(formula=formula, zoom_exponent=zoom_exponent, seed=SEED, **centre)
    (HERE / "c3_centres.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
