"""ctypes access to the test-only CPU build of the field kernel code
(fractal-audio-viz_amd/_build/libfav_host_check.so)."""
import ctypes

import numpy as np

import native

HOST_CHECK_SO = native.ROOT / "fractal-audio-viz_amd" / "_build" / "libfav_host_check.so"
# The same code built with a five-block derivative log (replays on a full log).
HOST_CHECK_LOG5_SO = native.ROOT / "fractal-audio-viz_amd" / "_build" / "libfav_host_check_log5.so"


def load(path=HOST_CHECK_SO):
    lib = ctypes.CDLL(str(path))
    lib.fav_host_check_render.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.POINTER(native.RenderOptions), native.c_float_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(native.RenderPlanes),
        native.c_uint64_p, ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_render.restype = ctypes.c_int
    return lib


def render(lib, meta):
    opts = native.RenderOptions.defaults(backend=2, **meta["options"])
    w, h = meta["width"], meta["height"]
    out = np.empty(w * h, dtype=np.float32)
    planes = native.PlaneArrays(w, h)
    stats = np.zeros(39, dtype=np.uint64)
    err = ctypes.create_string_buffer(512)
    status = lib.fav_host_check_render(
        meta["x"].encode(), meta["y"].encode(), meta["ref_zoom"].encode(), meta["max_iter"], meta["bits"],
        meta["image_series_order"], meta["clone_zoom"].encode() if meta["clone_zoom"] else None,
        ctypes.byref(opts), out.ctypes.data_as(native.c_float_p), w, h, meta["zoom"].encode(),
        meta["series_order"], meta["series_block"], ctypes.byref(planes.struct),
        stats.ctypes.data_as(native.c_uint64_p), err, 512)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out.reshape(h, w), planes, stats


def bind_kfp(lib):
    lib.fav_host_check_kfp.argtypes = [
        native.c_float_p, native.c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(native.KfpOptions), native.c_uint8_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_kfp.restype = ctypes.c_int
    return lib


def colourise(lib, field, max_iter, options, lut):
    """The colour kernel's per-pixel code run on the CPU."""
    h, w = field.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    f = np.ascontiguousarray(field, dtype=np.float32)
    l = np.ascontiguousarray(lut, dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    status = lib.fav_host_check_kfp(f.ctypes.data_as(native.c_float_p), out.ctypes.data_as(native.c_uint8_p),
                                    w, h, max_iter, ctypes.byref(options), l.ctypes.data_as(native.c_uint8_p),
                                    l.shape[0], err, 256)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out


def bind_direct(lib):
    lib.fav_host_check_direct.argtypes = [
        native.c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        ctypes.c_uint32, ctypes.POINTER(native.RenderPlanes), ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_direct.restype = ctypes.c_int
    return lib


def render_direct(lib, case):
    """The direct kernel's per-pixel code run on the CPU."""
    w, h = case["width"], case["height"]
    out = np.full(w * h, -7.0, dtype=np.float32)
    o = case.get("options", {})
    planes = native.PlaneArrays(w, h) if case["planes"] else None
    err = ctypes.create_string_buffer(256)
    status = lib.fav_host_check_direct(
        out.ctypes.data_as(native.c_float_p), w, h, case["zoom"].encode(), case["x"].encode(), case["y"].encode(),
        case["max_iter"], case["formula"], case.get("julia_re", 0.0), case.get("julia_im", 0.0),
        o.get("output_bias", 0.0), o.get("escape_radius_mode", 0), o.get("coordinate_mode", 0),
        o.get("plane_flags", 0), ctypes.byref(planes.struct) if planes is not None else None, err, 256)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out.reshape(h, w), planes


def bind_alt(lib):
    lib.fav_host_check_alt.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(native.RenderOptions), native.c_float_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_alt.restype = ctypes.c_int
    return lib


def render_alt(lib, case):
    """The alternate kernel's per-pixel code run on the CPU."""
    m = case
    w, h = m["width"], m["height"]
    out = np.full(w * h, -7.0, dtype=np.float32)
    opts = native.RenderOptions.defaults(backend=2, **m["options"])
    err = ctypes.create_string_buffer(512)
    status = lib.fav_host_check_alt(
        m["x"].encode(), m["y"].encode(), m["ref_zoom"].encode(), m["max_iter"], m["bits"], m["formula"],
        m["julia"][0].encode(), m["julia"][1].encode(), m["clone_zoom"].encode() if m["clone_zoom"] else None,
        ctypes.byref(opts), out.ctypes.data_as(native.c_float_p), w, h, m["zoom"].encode(), err, 512)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out.reshape(h, w)


def bind_fx(lib):
    lib.fav_host_check_fx.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(native.RenderOptions), native.c_float_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_fx.restype = ctypes.c_int
    lib.fav_host_check_fx_alt.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(native.RenderOptions), native.c_float_p, ctypes.c_void_p,
        native.c_float_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
        ctypes.c_int]
    lib.fav_host_check_fx_alt.restype = ctypes.c_int
    return lib


def render_fx(lib, meta):
    """The Mandelbrot floatexp path's per-pixel code on the CPU: (field,
    counts), the counts as the GPU entry writes them."""
    opts = native.RenderOptions.defaults(backend=2, **meta["options"])
    w, h = meta["width"], meta["height"]
    field = np.empty(w * h, dtype=np.float32)
    counts = np.empty(w * h, dtype=np.int32)
    err = ctypes.create_string_buffer(512)
    status = lib.fav_host_check_fx(meta["x"].encode(), meta["y"].encode(), meta["ref_zoom"].encode(),
                                   meta["max_iter"], meta["bits"], ctypes.byref(opts),
                                   field.ctypes.data_as(native.c_float_p), counts.ctypes.data, w, h,
                                   meta["zoom"].encode(), meta["series_order"], meta["series_block"], err, 512)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return field, counts


def render_fx_alt(lib, case):
    """The alternate-formula floatexp pixel and the FP64 pixel on the CPU:
    (fx field, fx counts, FP64 field, FP64 counts)."""
    opts = native.RenderOptions.defaults(backend=2, **case.get("options", {}))
    w, h = case["width"], case["height"]
    out = [np.empty(w * h, dtype=np.float32), np.empty(w * h, dtype=np.int32),
           np.empty(w * h, dtype=np.float32), np.empty(w * h, dtype=np.int32)]
    err = ctypes.create_string_buffer(512)
    status = lib.fav_host_check_fx_alt(
        case["x"].encode(), case["y"].encode(), case.get("ref_zoom", case["zoom"]).encode(), case["max_iter"],
        case["bits"], case["formula"], str(case["julia"][0]).encode(), str(case["julia"][1]).encode(),
        ctypes.byref(opts), out[0].ctypes.data_as(native.c_float_p), out[1].ctypes.data,
        out[2].ctypes.data_as(native.c_float_p), out[3].ctypes.data, w, h, case["zoom"].encode(), err, 512)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return tuple(out)
