"""Raw KFP atlas colour pass (renderer.h:663-711) through any library that
speaks the reference's C ABI, and the golden cases' inputs."""
import ctypes
import json
from pathlib import Path

import numpy as np

import native
from atlas_cases import CASES, case_fields
from kfp_cases import KALLES_DEFAULT_STOPS, seeded_stops

GOLDEN = Path(__file__).resolve().parent / "golden"
c_float_p = ctypes.POINTER(ctypes.c_float)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)

ATLAS_ARGS = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_int,
              ctypes.c_int, c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
              ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
              ctypes.POINTER(native.KfpOptions), c_uint8_p, ctypes.c_int, ctypes.c_int]


def bind_atlas(lib, entry="fractal_atlas_colourise_kfp_raw"):
    fn = getattr(lib, entry)
    fn.argtypes = ATLAS_ARGS
    fn.restype = ctypes.c_int
    lib.fractal_last_error.restype = ctypes.c_char_p
    return fn


def case_names():
    return [c["name"] for c in CASES]


def case_inputs(oracle, case):
    parent, child = case_fields(case)
    stops = np.asarray(KALLES_DEFAULT_STOPS if case["stops"] == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    return parent, child, lut, native.KfpOptions.make(**case["options"])


def atlas_call(fn, case, parent, child, lut, options, threads=0):
    ow, oh = case["output"]
    pw, ph, pmax = case["parent"]
    cw, ch, cmax = case["child"] if child is not None else (0, 0, 0)
    out = np.zeros((oh, ow, 3), dtype=np.uint8)
    p = np.ascontiguousarray(parent, dtype=np.float32)
    c = np.ascontiguousarray(child, dtype=np.float32) if child is not None else None
    l = np.ascontiguousarray(lut, dtype=np.uint8)
    status = fn(p.ctypes.data_as(c_float_p), pw, ph, pmax, c.ctypes.data_as(c_float_p) if c is not None else None,
                cw, ch, cmax, out.ctypes.data_as(c_uint8_p), ow, oh, case["parent_zoom"], case["child_fraction"],
                case["child_zoom"], case["max_iter"], 0.0, 0.0, 0.0, 0.0, ctypes.byref(options),
                l.ctypes.data_as(c_uint8_p), l.shape[0], threads)
    return status, out


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        return {"rgb": data["rgb"], "meta": json.loads(str(data["meta"]))}
