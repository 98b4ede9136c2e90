"""Point-render fixtures (tests/golden/points_*.npz) and the calls that
replay them through the C oracle and a renderer.h-speaking library."""
import ctypes
import json

import numpy as np

import native
from golden_io import GOLDEN
from point_cases import CASES


def case_names():
    return [c["name"] for c in CASES]


def case_by_name(name):
    return next(c for c in CASES if c["name"] == name)


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        out = {k: data[k] for k in data.files}
    out["meta"] = json.loads(str(out["meta"]))
    return out


def render_oracle(lib, case, points, stats=True):
    """fo_render_points on the case's reference and points."""
    re, im, e = (np.ascontiguousarray(a) for a in points)
    opts = native.RenderOptions.defaults(**case["options"])
    ref = native.OracleReference(lib, case["x"], case["y"], case["ref_zoom"], case["max_iter"], case["bits"],
                                 case["image_series_order"], case["options"].get("escape_radius_mode", 0),
                                 case["options"].get("coordinate_mode", 0), formula=case.get("formula", 0),
                                 julia=tuple(case.get("julia", ("0", "0"))))
    out = np.empty(len(re), dtype=np.float32)
    counters = np.zeros(39, dtype=np.uint64)
    err = ctypes.create_string_buffer(512)
    status = lib.fo_render_points(
        out.ctypes.data_as(native.c_float_p), len(re), case["zoom"].encode(),
        re.ctypes.data_as(native.c_double_p), im.ctypes.data_as(native.c_double_p),
        e.ctypes.data_as(native.c_int32_p), ref.handle, case["max_iter"], case["series_order"],
        case["series_block"], ctypes.byref(opts), counters.ctypes.data_as(native.c_uint64_p) if stats else None,
        0, err, 512)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out, counters


def render_abi(lib, case, points, backend=2, stats=True, handle=None):
    """fractal_render_points through the library; returns (field, stats)."""
    re, im, e = (np.ascontiguousarray(a) for a in points)
    opts = native.RenderOptions.defaults(backend=backend, **case["options"])
    ref = None
    if handle is None:
        ref = native.AbiReference(lib, case["x"], case["y"], case["ref_zoom"], case["max_iter"], case["bits"],
                                  case["image_series_order"], opts, formula=case.get("formula", 0),
                                  julia=tuple(case.get("julia", ("0", "0"))))
        handle = ref.handle
    out = np.empty(len(re), dtype=np.float32)
    if stats:
        lib.fractal_set_stats_enabled(1)
    status = lib.fractal_render_points(
        out.ctypes.data_as(native.c_float_p), len(re), case["zoom"].encode(),
        re.ctypes.data_as(native.c_double_p), im.ctypes.data_as(native.c_double_p),
        e.ctypes.data_as(native.c_int32_p), ctypes.c_void_p(handle), case["max_iter"], 0,
        case["series_order"], case["series_block"], ctypes.byref(opts))
    counters = None
    if stats:
        counters = np.zeros(39, dtype=np.uint64)
        lib.fractal_get_last_stats_ex(counters.ctypes.data_as(native.c_uint64_p), 39)
        lib.fractal_set_stats_enabled(0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out, counters
