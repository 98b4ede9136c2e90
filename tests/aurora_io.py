"""Inputs and calls of the ordinary-palette colouriser and the crop / atlas
compositors (renderer.h:400-449, 713-835, 943-1068) through any
renderer.h-speaking library: the reference's renderer.cpp (fixture
generation) or the HIP library (parity tests).  Inputs are regenerated from
each case's seed; the fixtures keep a checksum of them and the outputs."""
import ctypes
import hashlib
import json

import numpy as np

from golden_io import GOLDEN

F = ctypes.POINTER(ctypes.c_float)
U8 = ctypes.POINTER(ctypes.c_uint8)
I, D = ctypes.c_int, ctypes.c_double

SIGNATURES = {
    "fractal_colourise": [F, U8, I, I, I, D, D, D, D, I],
    "fractal_colourise_accents": [F, U8, I, I, I, D, D, D, D, U8, I, I, I, I],
    "fractal_apply_aurora_accents": [U8, I, I, U8, D, I],
    "fractal_crop_rgb": [U8, I, I, U8, I, I, D, I],
    "fractal_atlas_composite_rgb": [U8, I, I, U8, I, I, U8, I, I, D, D, D, I, I],
    "fractal_crop_field": [F, I, I, F, I, I, D, I],
    "fractal_atlas_field": [F, I, I, I, F, I, I, I, F, I, I, D, D, I, I],
    "fractal_atlas_field_ex": [F, I, I, I, F, I, I, I, F, I, I, D, D, D, D, D, D, I, I, I],
    "fractal_crop_colourise": [F, I, I, U8, I, I, D, I, D, D, D, D, I],
    "fractal_crop_colourise_interior": [F, I, I, U8, I, I, D, I, D, D, D, D, I, I, I, I],
    "fractal_crop_colourise_accents": [F, I, I, U8, I, I, D, I, D, D, D, D, U8, I, I, I, I],
    "fractal_atlas_colourise": [F, I, I, I, F, I, I, I, U8, I, I, D, D, I, D, D, D, D, I],
    "fractal_atlas_colourise_interior": [F, I, I, I, F, I, I, I, U8, I, I, D, D, I, D, D, D, D, I, I, I, I],
    "fractal_atlas_colourise_accents": [F, I, I, I, F, I, I, I, U8, I, I, D, D, I, D, D, D, D, U8, I, I, I, I],
}


def bind(lib):
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def field(seed, width, height, max_iter, bias=0.0):
    """A smooth escape-time-like field: bands in (0, 0.8 max_iter), interior
    blobs at max_iter, a few NaN pixels; minus `bias` (centred storage)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    phase = rng.uniform(0, 6.28, size=3)
    v = 0.4 * max_iter * (1.0 + np.sin(x / 7.3 + phase[0]) * np.cos(y / 5.1 + phase[1]))
    v += rng.uniform(0, 3.0, size=v.shape)
    cx, cy = rng.uniform(0.2, 0.8) * width, rng.uniform(0.2, 0.8) * height
    v[(x - cx) ** 2 + (y - cy) ** 2 < (0.15 * min(width, height)) ** 2] = max_iter
    v[rng.random(v.shape) < 0.004] = max_iter
    out = (v - bias).astype(np.float32)
    out[rng.random(out.shape) < 0.002] = np.nan
    return out


def rgb(seed, width, height):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    base = np.stack([(x * 5 + y * 3) % 256, (x * 2 + 40) % 256, (y * 7 + 90) % 256], axis=-1)
    return np.ascontiguousarray((base + rng.integers(0, 30, size=base.shape)) % 256, dtype=np.uint8)


def _f(a):
    return None if a is None else a.ctypes.data_as(F)


def _u(a):
    return None if a is None else a.ctypes.data_as(U8)


def inputs(case):
    """The case's source arrays, regenerated from its seed."""
    k = case["kind"]
    s = case["seed"]
    out = {}
    if k in ("colourise", "crop_colourise", "crop_field"):
        out["source"] = field(s, case["sw"], case["sh"], case["max_iter"])
    if k in ("atlas_field", "atlas_colourise"):
        out["parent"] = field(s, case["pw"], case["ph"], case["pmax"], case.get("pbias", 0.0))
        if case.get("child", True):
            out["child"] = field(s + 1, case["cw"], case["ch"], case["cmax"], case.get("cbias", 0.0))
    if k == "accents":
        out["rgb"] = rgb(s, case["w"], case["h"])
    if k == "crop_rgb":
        out["source"] = rgb(s, case["sw"], case["sh"])
    if k == "atlas_rgb":
        out["parent"] = rgb(s, case["pw"], case["ph"])
        if case.get("child", True):
            out["child"] = rgb(s + 1, case["cw"], case["ch"])
    if "accents" in case:
        out["accent_table"] = np.asarray(case["accents"], dtype=np.uint8)
    return out


def checksum(arrays):
    h = hashlib.sha256()
    for key in sorted(arrays):
        h.update(key.encode())
        h.update(np.ascontiguousarray(arrays[key]).tobytes())
    return h.hexdigest()


def run(lib, case):
    """Calls the case's function; returns (status, output array)."""
    a = inputs(case)
    k, fn = case["kind"], case["fn"]
    c = case
    acc = _u(a.get("accent_table"))
    interior = c.get("interior", (0, 0, 0))
    if k == "colourise":
        w, h = c["sw"], c["sh"]
        out = np.zeros((h, w, 3), dtype=np.uint8)
        args = [_f(a["source"]), _u(out), w, h, c["max_iter"], c["phase"], c["vocal"], c["instrumental"], c["pitch"]]
        if fn.endswith("accents"):
            args += [acc, *interior]
        status = getattr(lib, fn)(*args, 0)
    elif k == "accents":
        out = a["rgb"].copy()
        status = lib.fractal_apply_aurora_accents(_u(out), c["w"], c["h"], acc, c["pitch"], 0)
    elif k == "crop_rgb":
        out = np.zeros((c["oh"], c["ow"], 3), dtype=np.uint8)
        status = lib.fractal_crop_rgb(_u(a["source"]), c["sw"], c["sh"], _u(out), c["ow"], c["oh"], c["zoom"], 0)
    elif k == "atlas_rgb":
        out = np.zeros((c["oh"], c["ow"], 3), dtype=np.uint8)
        child = a.get("child")
        status = lib.fractal_atlas_composite_rgb(
            _u(a["parent"]), c["pw"], c["ph"], _u(child), c["cw"] if child is not None else 0,
            c["ch"] if child is not None else 0, _u(out), c["ow"], c["oh"], c["parent_zoom"], c["fraction"],
            c["child_zoom"], c["feather"], 0)
    elif k == "crop_field":
        out = np.zeros((c["oh"], c["ow"]), dtype=np.float32)
        status = lib.fractal_crop_field(_f(a["source"]), c["sw"], c["sh"], _f(out), c["ow"], c["oh"], c["zoom"], 0)
    elif k == "atlas_field":
        out = np.zeros((c["oh"], c["ow"]), dtype=np.float32)
        child = a.get("child")
        head = [_f(a["parent"]), c["pw"], c["ph"], c["pmax"], _f(child), c["cw"] if child is not None else 0,
                c["ch"] if child is not None else 0, c["cmax"] if child is not None else 0, _f(out), c["ow"], c["oh"],
                c["parent_zoom"], c["fraction"]]
        if fn.endswith("_ex"):
            status = lib.fractal_atlas_field_ex(*head, c["child_zoom"], c.get("pbias", 0.0), c.get("cbias", 0.0),
                                                c.get("obias", 0.0), c["palette_max_iter"], c["feather"], 0)
        else:
            status = lib.fractal_atlas_field(*head, c["palette_max_iter"], 0)
    elif k == "crop_colourise":
        out = np.zeros((c["oh"], c["ow"], 3), dtype=np.uint8)
        args = [_f(a["source"]), c["sw"], c["sh"], _u(out), c["ow"], c["oh"], c["zoom"], c["max_iter"], c["phase"],
                c["vocal"], c["instrumental"], c["pitch"]]
        if fn.endswith("interior"):
            args += list(interior)
        elif fn.endswith("accents"):
            args += [acc, *interior]
        status = getattr(lib, fn)(*args, 0)
    elif k == "atlas_colourise":
        out = np.zeros((c["oh"], c["ow"], 3), dtype=np.uint8)
        child = a.get("child")
        args = [_f(a["parent"]), c["pw"], c["ph"], c["pmax"], _f(child), c["cw"] if child is not None else 0,
                c["ch"] if child is not None else 0, c["cmax"] if child is not None else 0, _u(out), c["ow"], c["oh"],
                c["parent_zoom"], c["fraction"], c["palette_max_iter"], c["phase"], c["vocal"], c["instrumental"],
                c["pitch"]]
        if fn.endswith("interior"):
            args += list(interior)
        elif fn.endswith("accents"):
            args += [acc, *interior]
        status = getattr(lib, fn)(*args, 0)
    else:
        raise ValueError(k)
    return status, out


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        out = {k: data[k] for k in data.files}
    out["meta"] = json.loads(str(out["meta"]))
    return out
