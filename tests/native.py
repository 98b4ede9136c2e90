"""ctypes mirrors of the renderer.h ABI and loaders for the three libraries.

The struct mirrors follow renderer.h:43-170 (the same layout visualizer.py
declares at visualizer.py:824-944 and 986-1112).  Three shared objects speak
this ABI (or a subset):

* the HIP product, ``fractal-audio-viz_amd/libfractal_hip.so``;
* the C oracle, ``oracle/_build/libfav_oracle.so`` (test checker only);
* the reference's own renderer.cpp compiled from source into
  ``oracle/_ref/mandelbrot.so`` when the reference tree was available.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PRODUCT_SO = ROOT / "fractal-audio-viz_amd" / "libfractal_hip.so"
ORACLE_SO = ROOT / "oracle" / "_build" / "libfav_oracle.so"
REFERENCE_RELEASE_SO = ROOT / "oracle" / "_ref" / "mandelbrot_release.so"
REFERENCE_SO = ROOT / "oracle" / "_ref" / "mandelbrot.so"

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_uint64_p = ctypes.POINTER(ctypes.c_uint64)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)
c_int32_p = ctypes.POINTER(ctypes.c_int32)


class RenderOptions(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("version", ctypes.c_uint32),
        ("strict", ctypes.c_int32),
        ("allow_recovery", ctypes.c_int32),
        ("time_budget_ms", ctypes.c_int32),
        ("disable_bla", ctypes.c_int32),
        ("disable_cycle", ctypes.c_int32),
        ("strict_cycle", ctypes.c_int32),
        ("series_min_terms", ctypes.c_int32),
        ("series_max_terms", ctypes.c_int32),
        ("max_bla_length", ctypes.c_int32),
        ("max_linear_bla_length", ctypes.c_int32),
        ("backend", ctypes.c_int32),
        ("escape_radius_mode", ctypes.c_int32),
        ("coordinate_mode", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
        ("output_bias", ctypes.c_double),
    ]

    @classmethod
    def defaults(cls, **overrides) -> "RenderOptions":
        """renderer.cpp:4708-4727 defaults, then keyword overrides."""
        o = cls()
        o.struct_size = ctypes.sizeof(cls)
        o.version = 2
        o.strict = 1
        o.series_min_terms = 8
        o.series_max_terms = 32
        o.max_bla_length = 64
        o.max_linear_bla_length = 4096
        for key, value in overrides.items():
            setattr(o, key, value)
        return o


class RenderPlanes(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("version", ctypes.c_uint32),
        ("orbit_iteration", c_int64_p),
        ("phase", c_double_p),
        ("de_x", c_double_p),
        ("de_y", c_double_p),
        ("test1", c_double_p),
        ("test2", c_double_p),
    ]


class PlaneArrays:
    """Owns numpy storage for a FractalRenderPlanes struct."""

    NAMES = ("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")

    def __init__(self, width: int, height: int):
        n = width * height
        self.orbit_iteration = np.full(n, -7, dtype=np.int64)
        self.phase = np.full(n, -7.0)
        self.de_x = np.full(n, -7.0)
        self.de_y = np.full(n, -7.0)
        self.test1 = np.full(n, -7.0)
        self.test2 = np.full(n, -7.0)
        self.struct = RenderPlanes()
        self.struct.struct_size = ctypes.sizeof(RenderPlanes)
        self.struct.version = 1
        self.struct.orbit_iteration = self.orbit_iteration.ctypes.data_as(c_int64_p)
        for name in self.NAMES[1:]:
            setattr(self.struct, name, getattr(self, name).ctypes.data_as(c_double_p))

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name in self.NAMES}


def _bind_render_abi(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Argument types for the renderer.h entry points (renderer.h:172-373)."""
    lib.fractal_abi_version.restype = ctypes.c_int
    lib.fractal_last_error.restype = ctypes.c_char_p
    lib.fractal_backend_capabilities.restype = ctypes.c_int
    lib.fractal_render_options_default.argtypes = [ctypes.POINTER(RenderOptions)]
    lib.fractal_set_stats_enabled.argtypes = [ctypes.c_int]
    for name in ("fractal_get_last_stats", "fractal_get_last_stats_ex"):
        fn = getattr(lib, name)
        fn.argtypes = [c_uint64_p, ctypes.c_int]
        fn.restype = ctypes.c_int
    lib.fractal_create_reference_ex_options.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
        ctypes.POINTER(RenderOptions)]
    lib.fractal_create_reference_ex_options.restype = ctypes.c_void_p
    for name in ("fractal_create_reference", "fractal_create_reference_reusable"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                       ctypes.c_int, ctypes.c_int]
        fn.restype = ctypes.c_void_p
    lib.fractal_create_reference_reusable_options.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(RenderOptions)]
    lib.fractal_create_reference_reusable_options.restype = ctypes.c_void_p
    lib.fractal_create_reference_ex.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p]
    lib.fractal_create_reference_ex.restype = ctypes.c_void_p
    lib.fractal_clone_reference.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.fractal_clone_reference.restype = ctypes.c_void_p
    lib.fractal_destroy_reference.argtypes = [ctypes.c_void_p]
    lib.fractal_destroy_reference.restype = None
    lib.fractal_get_reference_stats.argtypes = [ctypes.c_void_p, c_uint64_p, ctypes.c_int]
    lib.fractal_get_reference_stats.restype = ctypes.c_int
    lib.fractal_render_reference_ex_planes.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(RenderOptions), ctypes.POINTER(RenderPlanes)]
    lib.fractal_render_reference_ex_planes.restype = ctypes.c_int
    for name in ("fractal_render_reference_ex", "fractal_render_mandelbrot_reference_ex"):
        fn = getattr(lib, name)
        fn.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                       ctypes.POINTER(RenderOptions)]
        fn.restype = ctypes.c_int
    lib.render_fractal_ex.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
        ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, ctypes.POINTER(RenderOptions)]
    lib.render_fractal_ex.restype = ctypes.c_int
    if hasattr(lib, "render_fractal_ex_planes"):
        lib.render_fractal_ex_planes.argtypes = lib.render_fractal_ex.argtypes + [ctypes.POINTER(RenderPlanes)]
        lib.render_fractal_ex_planes.restype = ctypes.c_int
    if hasattr(lib, "render_mandelbrot_ex"):
        lib.render_mandelbrot_ex.argtypes = [
            c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(RenderOptions)]
        lib.render_mandelbrot_ex.restype = ctypes.c_int
    for name in ("fractal_render_points", "fractal_render_points_ex_planes"):
        if hasattr(lib, name):  # renderer.h:287-321
            fn = getattr(lib, name)
            fn.argtypes = [c_float_p, ctypes.c_int, ctypes.c_char_p, c_double_p, c_double_p,
                           c_int32_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.POINTER(RenderOptions)]
            if name.endswith("planes"):
                fn.argtypes = fn.argtypes + [ctypes.POINTER(RenderPlanes)]
            fn.restype = ctypes.c_int
    if hasattr(lib, "render_mandelbrot"):
        lib.render_mandelbrot.argtypes = [
            c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.render_mandelbrot.restype = ctypes.c_int
    return lib


def direct_options(case, backend):
    """FractalRenderOptions of a direct-render case."""
    o = case.get("options", {})
    opts = RenderOptions.defaults(backend=backend, output_bias=o.get("output_bias", 0.0),
                                  escape_radius_mode=o.get("escape_radius_mode", 0),
                                  coordinate_mode=o.get("coordinate_mode", 0))
    opts.reserved[0] = o.get("plane_flags", 0)
    return opts


def render_direct_abi(lib, case, backend=2):
    """render_fractal_ex[_planes] with use_perturbation = 0."""
    w, h = case["width"], case["height"]
    out = np.full(w * h, -7.0, dtype=np.float32)
    opts = direct_options(case, backend)
    planes = PlaneArrays(w, h) if case["planes"] else None
    args = [out.ctypes.data_as(c_float_p), w, h, case["zoom"].encode(), case["x"].encode(), case["y"].encode(),
            case["max_iter"], 128, 0, 0, case["formula"], case.get("julia_re", 0.0), case.get("julia_im", 0.0),
            ctypes.byref(opts)]
    if planes is not None:
        status = lib.render_fractal_ex_planes(*args, ctypes.byref(planes.struct))
    else:
        status = lib.render_fractal_ex(*args)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out.reshape(h, w), planes


def render_direct_oracle(lib, case, threads=0):
    lib.fo_render_direct.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        ctypes.c_uint32, ctypes.POINTER(RenderPlanes), ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    w, h = case["width"], case["height"]
    out = np.full(w * h, -7.0, dtype=np.float32)
    o = case.get("options", {})
    planes = PlaneArrays(w, h) if case["planes"] else None
    err = ctypes.create_string_buffer(256)
    status = lib.fo_render_direct(
        out.ctypes.data_as(c_float_p), w, h, case["zoom"].encode(), case["x"].encode(), case["y"].encode(),
        case["max_iter"], case["formula"], case.get("julia_re", 0.0), case.get("julia_im", 0.0),
        o.get("output_bias", 0.0), o.get("escape_radius_mode", 0), o.get("coordinate_mode", 0),
        o.get("plane_flags", 0), ctypes.byref(planes.struct) if planes is not None else None, threads, err, 256)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out.reshape(h, w), planes


def load_reference(release=False):
    """The reference renderer.cpp built from source, or None if absent.
    release=True: the build with the reference Makefile's own flags (FMA
    contraction on), used only as the CPU timing baseline."""
    path = REFERENCE_RELEASE_SO if release else REFERENCE_SO
    if not path.exists():
        return None
    return _bind_render_abi(ctypes.CDLL(str(path)))


def load_oracle():
    lib = ctypes.CDLL(str(ORACLE_SO))
    lib.fo_reference_create.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.fo_reference_create.restype = ctypes.c_void_p
    lib.fo_reference_clone.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.fo_reference_clone.restype = ctypes.c_void_p
    lib.fo_reference_destroy.argtypes = [ctypes.c_void_p]
    lib.fo_reference_info.argtypes = [ctypes.c_void_p, c_int64_p, ctypes.c_int]
    lib.fo_render.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.POINTER(RenderOptions), ctypes.POINTER(RenderPlanes),
        c_uint64_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.fo_render.restype = ctypes.c_int
    lib.fo_render_crop.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(RenderOptions),
        ctypes.POINTER(RenderPlanes), c_uint64_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.fo_render_crop.restype = ctypes.c_int
    lib.fo_render_points.argtypes = [
        c_float_p, ctypes.c_int, ctypes.c_char_p, c_double_p, c_double_p, c_int32_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(RenderOptions), c_uint64_p, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_int]
    lib.fo_render_points.restype = ctypes.c_int
    return lib


def load_product():
    """The HIP library; raises if it has not been built."""
    if not PRODUCT_SO.exists():
        raise FileNotFoundError(f"{PRODUCT_SO} is not built; run __graft_entry__.build()")
    return bind_kfp(_bind_render_abi(ctypes.CDLL(str(PRODUCT_SO))))


def _b(text) -> bytes:
    return text.encode() if isinstance(text, str) else text


class OracleReference:
    """A fo_reference owned by Python."""

    def __init__(self, lib, x, y, zoom, max_iter, precision_bits, series_order=8,
                 escape_radius_mode=0, coordinate_mode=0, handle=None, formula=0, julia=("0", "0"),
                 bla_epsilon_exponent=-38):
        self.lib = lib
        if handle is None:
            err = ctypes.create_string_buffer(512)
            lib.fo_reference_create_tuned.argtypes = [
                ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                ctypes.c_char_p, ctypes.c_int]
            lib.fo_reference_create_tuned.restype = ctypes.c_void_p
            handle = lib.fo_reference_create_tuned(_b(x), _b(y), _b(zoom), max_iter, precision_bits,
                                                    series_order, escape_radius_mode, coordinate_mode,
                                                    formula, _b(julia[0]), _b(julia[1]), bla_epsilon_exponent,
                                                    err, 512)
            if not handle:
                raise RuntimeError(err.value.decode())
        self.handle = handle

    def clone(self, zoom):
        err = ctypes.create_string_buffer(512)
        handle = self.lib.fo_reference_clone(self.handle, _b(zoom), err, 512)
        if not handle:
            raise RuntimeError(err.value.decode())
        return OracleReference(self.lib, None, None, None, 0, 0, handle=handle)

    def info(self):
        values = np.zeros(9, dtype=np.int64)
        self.lib.fo_reference_info(self.handle, values.ctypes.data_as(c_int64_p), 9)
        keys = ("orbit_size", "map_end", "series_enabled", "series_iteration", "series_order",
                "cubic_levels", "linear_levels", "deep_linear_levels", "cubic_base")
        return dict(zip(keys, values.tolist()))

    def render(self, width, height, zoom, max_iter, series_order=8, series_block=32,
               options=None, planes=False, stats=False, threads=0):
        out = np.empty(width * height, dtype=np.float32)
        opts = options if options is not None else RenderOptions.defaults()
        plane_arrays = PlaneArrays(width, height) if planes else None
        stat_values = np.zeros(39, dtype=np.uint64) if stats else None
        err = ctypes.create_string_buffer(512)
        status = self.lib.fo_render(
            out.ctypes.data_as(c_float_p), width, height, _b(zoom), self.handle, max_iter,
            series_order, series_block, ctypes.byref(opts),
            ctypes.byref(plane_arrays.struct) if plane_arrays else None,
            stat_values.ctypes.data_as(c_uint64_p) if stats else None, threads, err, 512)
        if status != 0:
            raise RuntimeError(err.value.decode())
        return out.reshape(height, width), plane_arrays, stat_values

    def render_crop(self, width, height, crop, zoom, max_iter, series_order=8, series_block=32,
                    options=None, planes=False, stats=False, threads=0):
        """The pixels of crop = (x, y, w, h) of the width x height field, as
        an h x w array equal to that block of the full render."""
        x, y, w, h = crop
        out = np.empty(w * h, dtype=np.float32)
        opts = options if options is not None else RenderOptions.defaults()
        plane_arrays = PlaneArrays(w, h) if planes else None
        stat_values = np.zeros(39, dtype=np.uint64) if stats else None
        err = ctypes.create_string_buffer(512)
        status = self.lib.fo_render_crop(
            out.ctypes.data_as(c_float_p), width, height, x, y, w, h, _b(zoom), self.handle, max_iter,
            series_order, series_block, ctypes.byref(opts),
            ctypes.byref(plane_arrays.struct) if plane_arrays else None,
            stat_values.ctypes.data_as(c_uint64_p) if stats else None, threads, err, 512)
        if status != 0:
            raise RuntimeError(err.value.decode())
        return out.reshape(h, w), plane_arrays, stat_values

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.fo_reference_destroy(self.handle)
            self.handle = None


class AbiReference:
    """A reference handle created through a renderer.h-speaking library."""

    def __init__(self, lib, x, y, zoom, max_iter, precision_bits, series_order=8,
                 options=None, formula=0, julia=("0", "0"), bla_epsilon_exponent=None):
        self.lib = lib
        opts = options if options is not None else RenderOptions.defaults()
        if bla_epsilon_exponent is None:
            handle = lib.fractal_create_reference_ex_options(
                _b(x), _b(y), _b(zoom), max_iter, precision_bits, series_order, formula,
                _b(julia[0]), _b(julia[1]), ctypes.byref(opts))
        else:  # the HIP library's extension entry
            lib.fractal_hip_create_reference_tuned.argtypes = [
                ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(RenderOptions), ctypes.c_int]
            lib.fractal_hip_create_reference_tuned.restype = ctypes.c_void_p
            handle = lib.fractal_hip_create_reference_tuned(
                _b(x), _b(y), _b(zoom), max_iter, precision_bits, series_order, formula,
                _b(julia[0]), _b(julia[1]), ctypes.byref(opts), int(bla_epsilon_exponent))
        if not handle:
            raise RuntimeError(lib.fractal_last_error().decode())
        self.handle = handle

    def render(self, width, height, zoom, max_iter, series_order=8, series_block=32,
               options=None, planes=False, stats=False, threads=0):
        out = np.empty(width * height, dtype=np.float32)
        opts = options if options is not None else RenderOptions.defaults()
        plane_arrays = PlaneArrays(width, height) if planes else None
        if stats:
            self.lib.fractal_set_stats_enabled(1)
        status = self.lib.fractal_render_reference_ex_planes(
            out.ctypes.data_as(c_float_p), width, height, _b(zoom), ctypes.c_void_p(self.handle),
            max_iter, threads, series_order, series_block, ctypes.byref(opts),
            ctypes.byref(plane_arrays.struct) if plane_arrays else None)
        stat_values = None
        if stats:
            stat_values = np.zeros(39, dtype=np.uint64)
            self.lib.fractal_get_last_stats_ex(stat_values.ctypes.data_as(c_uint64_p), 39)
            self.lib.fractal_set_stats_enabled(0)
        if status != 0:
            raise RuntimeError(self.lib.fractal_last_error().decode())
        return out.reshape(height, width), plane_arrays, stat_values

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.fractal_destroy_reference(ctypes.c_void_p(self.handle))
            self.handle = None


def escape_counts(field: np.ndarray, planes: PlaneArrays, max_iter: int) -> np.ndarray:
    """n_p = orbit_iteration + 1 for escaped pixels, max_iter for interior
    and unresolved (NaN) pixels; orbit_iteration is Kalles' antal = n - 1
    (renderer.cpp:8098-8106)."""
    counts = planes.orbit_iteration.astype(np.int64) + 1
    interior = planes.orbit_iteration >= max_iter
    counts[interior] = max_iter
    counts[np.isnan(field.reshape(-1))] = max_iter
    return counts.reshape(field.shape)


class KfpOptions(ctypes.Structure):
    """FractalKfpOptions, renderer.h:75-119 (version 3)."""

    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
        ("iter_div", ctypes.c_double), ("color_offset", ctypes.c_double), ("ratio", ctypes.c_double),
        ("color_method", ctypes.c_int32), ("smooth_method", ctypes.c_int32), ("smooth", ctypes.c_int32),
        ("flat", ctypes.c_int32), ("inverse_transition", ctypes.c_int32),
        ("phase_color_strength", ctypes.c_double), ("multi_color", ctypes.c_int32),
        ("blend_multi_color", ctypes.c_int32), ("multi_color_count", ctypes.c_uint32),
        ("multi_color_period", ctypes.c_double * 256), ("multi_color_start", ctypes.c_int32 * 256),
        ("multi_color_type", ctypes.c_int32 * 256), ("power", ctypes.c_double), ("slopes", ctypes.c_int32),
        ("slope_power", ctypes.c_double), ("slope_ratio", ctypes.c_double), ("slope_angle", ctypes.c_double),
        ("differences", ctypes.c_int32), ("interior_color", ctypes.c_int32 * 3), ("field_bias", ctypes.c_double),
        ("bailout_radius_preset", ctypes.c_int32), ("bailout_radius_custom", ctypes.c_double),
        ("bailout_norm_preset", ctypes.c_int32), ("bailout_norm_custom", ctypes.c_double),
        ("texture_enabled", ctypes.c_int32), ("texture_merge", ctypes.c_double),
        ("texture_power", ctypes.c_double), ("texture_ratio", ctypes.c_double),
        ("texture_resize", ctypes.c_int32), ("use_opengl", ctypes.c_int32), ("use_srgb", ctypes.c_int32),
        ("show_glitches", ctypes.c_int32),
    ]

    @classmethod
    def make(cls, **kw) -> "KfpOptions":
        """Kalles defaults (visualizer.py:1163-1215 NativeKfpOptions.from_profile
        field meanings), overridden by keywords; multi_colors is a list of
        (period, type) pairs."""
        o = cls()
        o.struct_size = ctypes.sizeof(cls)
        o.version = 3
        o.iter_div = 1.0
        o.ratio = 360.0
        o.smooth = 1
        o.power = 2.0
        o.slope_power = 50.0
        o.slope_ratio = 20.0
        o.slope_angle = 45.0
        o.bailout_radius_custom = 2.0
        o.bailout_norm_custom = 2.0
        o.texture_merge = 1.0
        o.texture_power = 200.0
        o.texture_ratio = 100.0
        o.show_glitches = 1
        multi = kw.pop("multi_colors", None)
        interior = kw.pop("interior_color", (0, 0, 0))
        for key, value in kw.items():
            setattr(o, key, value)
        o.interior_color[0], o.interior_color[1], o.interior_color[2] = interior
        if multi:
            o.multi_color = 1
            o.multi_color_count = len(multi)
            for i, (period, kind) in enumerate(multi):
                o.multi_color_period[i] = period
                o.multi_color_type[i] = kind
        return o

    def as_dict(self) -> dict:
        out = {}
        for name, ctype in self._fields_:
            value = getattr(self, name)
            if hasattr(value, "__len__"):
                value = list(value)
            out[name] = value
        return out

    @classmethod
    def from_dict(cls, data: dict) -> "KfpOptions":
        o = cls()
        for name, _ in cls._fields_:
            value = data[name]
            if isinstance(value, list):
                arr = getattr(o, name)
                for i, v in enumerate(value):
                    arr[i] = v
            else:
                setattr(o, name, value)
        return o


def bind_kfp(lib):
    """fractal_colourise_kfp[_planes] argument types (renderer.h:452-498)."""
    lib.fractal_colourise_kfp.argtypes = [
        c_float_p, c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, ctypes.POINTER(KfpOptions), c_uint8_p, ctypes.c_int, ctypes.c_int]
    lib.fractal_colourise_kfp.restype = ctypes.c_int
    lib.fractal_colourise_kfp_planes.argtypes = [
        c_float_p, c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, ctypes.POINTER(KfpOptions), ctypes.c_void_p, c_uint8_p, ctypes.c_int,
        ctypes.c_int]
    lib.fractal_colourise_kfp_planes.restype = ctypes.c_int
    for name in ("fractal_colourise_kfp_opencl", "fractal_colourise_kfp_precise"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = lib.fractal_colourise_kfp.argtypes
            getattr(lib, name).restype = ctypes.c_int
    if hasattr(lib, "fractal_hip_colourise_kfp_device"):
        lib.fractal_hip_colourise_kfp_device.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KfpOptions),
            c_uint8_p, ctypes.c_int]
        lib.fractal_hip_colourise_kfp_device.restype = ctypes.c_int
    if hasattr(lib, "fractal_hip_colourise_kfp_device_noalloc"):
        lib.fractal_hip_colourise_kfp_device_noalloc.argtypes = lib.fractal_hip_colourise_kfp_device.argtypes
        lib.fractal_hip_colourise_kfp_device_noalloc.restype = ctypes.c_int
    return lib


def colourise_abi(lib, field, max_iter, options, lut, threads=0, entry="fractal_colourise_kfp"):
    h, w = field.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    f = np.ascontiguousarray(field, dtype=np.float32)
    l = np.ascontiguousarray(lut, dtype=np.uint8)
    status = getattr(lib, entry)(f.ctypes.data_as(c_float_p), out.ctypes.data_as(c_uint8_p), w, h,
                                 max_iter, 0.0, 0.0, 0.0, 0.0, ctypes.byref(options),
                                 l.ctypes.data_as(c_uint8_p), l.shape[0], threads)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def colourise_oracle(lib, field, max_iter, options, lut, threads=0):
    lib.fo_colourise_kfp.argtypes = [
        c_float_p, c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(KfpOptions),
        ctypes.c_void_p, c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    h, w = field.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    f = np.ascontiguousarray(field, dtype=np.float32)
    l = np.ascontiguousarray(lut, dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    status = lib.fo_colourise_kfp(f.ctypes.data_as(c_float_p), out.ctypes.data_as(c_uint8_p), w, h, max_iter,
                                  ctypes.byref(options), None, l.ctypes.data_as(c_uint8_p), l.shape[0], threads,
                                  err, 256)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out


def palette_lut_oracle(lib, stops, size=1024):
    lib.fo_kfp_palette_lut.argtypes = [c_uint8_p, ctypes.c_int, c_uint8_p, ctypes.c_int]
    s = np.ascontiguousarray(stops, dtype=np.uint8)
    lut = np.zeros((size, 3), dtype=np.uint8)
    assert lib.fo_kfp_palette_lut(s.ctypes.data_as(c_uint8_p), s.shape[0], lut.ctypes.data_as(c_uint8_p), size) == 0
    return lut
