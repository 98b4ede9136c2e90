"""Kalles' 1024-entry cyclic palette table (the host side of the KFP colour
pass; visualizer.py:13491-13521 _kfp_palette_lut).

Entry i sits at position p = i*S/size among the S key colours; the blend
weight between key floor(p) and the next key (cyclically) is the sine ramp
sin((frac(p) - 1/2)*pi)/2 + 1/2, computed in float64 and truncated to a
byte after clipping to [0, 255].
"""
import numpy as np


def kfp_palette_lut(stops, size: int = 1024) -> np.ndarray:
    """(size, 3) uint8 table from S >= 2 RGB key colours."""
    keys = np.asarray(stops, dtype=np.float64).reshape(-1, 3)
    count = keys.shape[0]
    if count < 2:
        raise ValueError("KFP palette needs at least two colour stops")
    size = max(2, int(size))
    position = np.arange(size, dtype=np.float64) * float(count) / float(size)
    first = np.floor(position).astype(np.intp)
    second = (first + 1) % count
    ramp = np.sin((position - first.astype(np.float64) - 0.5) * np.pi) / 2.0 + 0.5
    blended = keys[first] * (1.0 - ramp[:, None]) + keys[second] * ramp[:, None]
    return np.clip(blended, 0.0, 255.0).astype(np.uint8)


def seeded_stops(count: int, seed: int) -> np.ndarray:
    """S uniform RGB key colours from numpy's default generator."""
    return np.random.default_rng(seed).integers(0, 256, (count, 3)).astype(np.uint8)
