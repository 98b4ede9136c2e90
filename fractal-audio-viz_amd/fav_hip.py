"""Python host binding of the HIP deep-zoom backend.

Mirrors the parts of visualizer.py's ctypes layer that drive the field path
(visualizer.py:824-944 NativeRenderOptions, 986-1112 planes, 1512-1560 the
loader's ABI check, 6250-6370 reference creation, 6529-6639 reference render)
so a caller written against the reference's binding finds the same names,
argument meaning and errors.  The library must be built (make -C
fractal-audio-viz_amd); there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "libfractal_hip.so"
ABI_VERSION = 10

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)
_c_uint64_p = ctypes.POINTER(ctypes.c_uint64)
_c_uint8_p = ctypes.POINTER(ctypes.c_uint8)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)


class NativeRenderOptions(ctypes.Structure):
    """FractalRenderOptions (renderer.h:43-67), same defaults as
    visualizer.py:840-944 except backend=2 (the HIP device)."""

    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
        ("strict", ctypes.c_int32), ("allow_recovery", ctypes.c_int32),
        ("time_budget_ms", ctypes.c_int32), ("disable_bla", ctypes.c_int32),
        ("disable_cycle", ctypes.c_int32), ("strict_cycle", ctypes.c_int32),
        ("series_min_terms", ctypes.c_int32), ("series_max_terms", ctypes.c_int32),
        ("max_bla_length", ctypes.c_int32), ("max_linear_bla_length", ctypes.c_int32),
        ("backend", ctypes.c_int32), ("escape_radius_mode", ctypes.c_int32),
        ("coordinate_mode", ctypes.c_int32), ("reserved", ctypes.c_int32 * 1),
        ("output_bias", ctypes.c_double),
    ]

    def __init__(self, *, strict=True, allow_recovery=False, time_budget_ms=0, disable_bla=False,
                 disable_cycle=False, strict_cycle=False, series_min_terms=8, series_max_terms=32,
                 max_bla_length=64, max_linear_bla_length=4096, backend=2, escape_radius_mode=0,
                 coordinate_mode=0, output_bias=0.0):
        super().__init__()
        self.struct_size = ctypes.sizeof(NativeRenderOptions)
        self.version = 2
        self.strict = int(bool(strict))
        self.allow_recovery = int(bool(allow_recovery))
        self.time_budget_ms = int(time_budget_ms)
        self.disable_bla = int(bool(disable_bla))
        self.disable_cycle = int(bool(disable_cycle))
        self.strict_cycle = int(bool(strict_cycle))
        self.series_min_terms = int(series_min_terms)
        self.series_max_terms = int(series_max_terms)
        self.max_bla_length = int(max_bla_length)
        self.max_linear_bla_length = int(max_linear_bla_length)
        if backend not in (0, 1, 2):
            raise ValueError("backend must be 0, 1 or 2")
        self.backend = int(backend)
        self.escape_radius_mode = int(escape_radius_mode)
        self.coordinate_mode = int(coordinate_mode)
        self.output_bias = float(output_bias)


class NativeRenderPlanes(ctypes.Structure):
    """FractalRenderPlanes (renderer.h:161-170)."""

    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
        ("orbit_iteration", _c_int64_p), ("phase", _c_double_p), ("de_x", _c_double_p),
        ("de_y", _c_double_p), ("test1", _c_double_p), ("test2", _c_double_p),
    ]


class NativeKfpOptions(ctypes.Structure):
    """FractalKfpOptions version 3 (renderer.h:75-119); the keyword
    defaults are Kalles' (visualizer.py:1163-1215 field meanings)."""

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

    def __init__(self, **kw):
        super().__init__()
        self.struct_size = ctypes.sizeof(NativeKfpOptions)
        self.version = 3
        defaults = dict(iter_div=1.0, ratio=360.0, smooth=1, power=2.0, slope_power=50.0, slope_ratio=20.0,
                        slope_angle=45.0, bailout_radius_custom=2.0, bailout_norm_custom=2.0, texture_merge=1.0,
                        texture_power=200.0, texture_ratio=100.0, show_glitches=1)
        defaults.update(kw)
        for key, value in defaults.items():
            setattr(self, key, value)

    @classmethod
    def kalles_default(cls) -> "NativeKfpOptions":
        """KALLES_DEFAULT_KFP's transfer (visualizer.py:430-444): method 7
        (field-difference distance), differences 3, slopes at 45 degrees."""
        return cls(iter_div=0.01, ratio=360.0, color_method=7, smooth_method=0, smooth=1, power=2.0, slopes=1,
                   slope_power=50.0, slope_ratio=20.0, slope_angle=45.0, differences=3)


_library = None


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Loads libfractal_hip.so and checks its ABI version (visualizer.py:1550)."""
    global _library
    if _library is not None and path is None:
        return _library
    target = Path(path or os.environ.get("FRACTAL_HIP_LIBRARY", LIBRARY_PATH))
    if not target.exists():
        raise FileNotFoundError(f"{target} is not built; run make -C {LIBRARY_PATH.parent}")
    lib = ctypes.CDLL(str(target))
    lib.fractal_abi_version.restype = ctypes.c_int
    if lib.fractal_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{target} speaks ABI {lib.fractal_abi_version()}, expected {ABI_VERSION}")
    lib.fractal_last_error.restype = ctypes.c_char_p
    lib.fractal_backend_capabilities.restype = ctypes.c_int
    lib.fractal_create_reference_ex_options.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(NativeRenderOptions)]
    lib.fractal_create_reference_ex_options.restype = ctypes.c_void_p
    lib.fractal_destroy_reference.argtypes = [ctypes.c_void_p]
    lib.fractal_destroy_reference.restype = None
    lib.fractal_render_reference_ex_planes.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(NativeRenderOptions),
        ctypes.POINTER(NativeRenderPlanes)]
    lib.fractal_render_reference_ex_planes.restype = ctypes.c_int
    lib.fractal_render_points.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_char_p, _c_double_p, _c_double_p, _c_int32_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(NativeRenderOptions)]
    lib.fractal_render_points.restype = ctypes.c_int
    lib.fractal_hip_render_reference_device.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.POINTER(NativeRenderOptions), ctypes.c_void_p]
    lib.fractal_hip_render_reference_device.restype = ctypes.c_int
    lib.fractal_hip_render_reference_floatexp.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(NativeRenderOptions), ctypes.c_int]
    lib.fractal_hip_render_reference_floatexp.restype = ctypes.c_int
    lib.fractal_hip_render_reference_counts.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(NativeRenderOptions)]
    lib.fractal_hip_render_reference_counts.restype = ctypes.c_int
    lib.fractal_hip_batch_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NativeRenderOptions)]
    lib.fractal_hip_batch_create.restype = ctypes.c_void_p
    lib.fractal_hip_batch_tile_count.argtypes = [ctypes.c_void_p]
    lib.fractal_hip_batch_tile_count.restype = ctypes.c_int
    lib.fractal_hip_batch_set_order.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    lib.fractal_hip_batch_set_order.restype = ctypes.c_int
    lib.fractal_hip_batch_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.fractal_hip_batch_set_precision.restype = ctypes.c_int
    lib.fractal_hip_batch_render.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.fractal_hip_batch_render.restype = ctypes.c_int
    lib.fractal_hip_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.fractal_hip_batch_destroy.restype = None
    lib.fractal_hip_render_segment.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NativeRenderOptions),
                                               ctypes.c_int, ctypes.c_int, _c_double_p]
    lib.fractal_hip_render_segment.restype = ctypes.c_int
    lib.fractal_hip_last_timing.argtypes = [_c_double_p, ctypes.c_int]
    lib.fractal_hip_last_timing.restype = ctypes.c_int
    lib.fractal_hip_reference_upload.argtypes = [ctypes.c_void_p]
    lib.fractal_hip_reference_upload.restype = ctypes.c_int
    colour_args = [_c_float_p, _c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(NativeKfpOptions), _c_uint8_p,
                   ctypes.c_int, ctypes.c_int]
    for name in ("fractal_colourise_kfp", "fractal_colourise_kfp_opencl", "fractal_colourise_kfp_precise"):
        getattr(lib, name).argtypes = colour_args
        getattr(lib, name).restype = ctypes.c_int
    lib.fractal_hip_colourise_kfp_device.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(NativeKfpOptions), _c_uint8_p, ctypes.c_int]
    lib.fractal_hip_colourise_kfp_device.restype = ctypes.c_int
    lib.fractal_hip_render_reference_scheduled.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(NativeRenderOptions), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.fractal_hip_render_reference_scheduled.restype = ctypes.c_int
    lib.fractal_hip_preview_features.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(NativeRenderOptions), ctypes.c_int]
    lib.fractal_hip_preview_features.restype = ctypes.c_int
    lib.fractal_hip_render_reference_learned.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(NativeRenderOptions), _c_float_p, ctypes.c_int, ctypes.c_void_p, _c_double_p]
    lib.fractal_hip_render_reference_learned.restype = ctypes.c_int
    lib.fractal_hip_cost_model_parameter_count.argtypes = []
    lib.fractal_hip_cost_model_parameter_count.restype = ctypes.c_int
    lib.fractal_hip_cost_model_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.fractal_hip_cost_model_forward.restype = ctypes.c_int
    lib.fractal_hip_cost_model_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_void_p]
    lib.fractal_hip_cost_model_gradient.restype = ctypes.c_int
    lib.fractal_hip_cost_model_train.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
    lib.fractal_hip_cost_model_train.restype = ctypes.c_int
    lib.fractal_atlas_colourise_kfp_raw.argtypes = [
        _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        _c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(NativeKfpOptions),
        _c_uint8_p, ctypes.c_int, ctypes.c_int]
    lib.fractal_atlas_colourise_kfp_raw.restype = ctypes.c_int
    lib.fractal_atlas_colourise_kfp_raw_precise.argtypes = lib.fractal_atlas_colourise_kfp_raw.argtypes
    lib.fractal_atlas_colourise_kfp_raw_precise.restype = ctypes.c_int
    # ordinary palette and per-frame compositors (renderer.h:400-449, 713-835, 943-1068)
    F, U8, I, D = _c_float_p, _c_uint8_p, ctypes.c_int, ctypes.c_double
    for name, args in {
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
    }.items():
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = ctypes.c_int
    lib.fractal_set_stats_enabled.argtypes = [ctypes.c_int]
    lib.fractal_get_last_stats_ex.argtypes = [_c_uint64_p, ctypes.c_int]
    lib.fractal_get_last_stats_ex.restype = ctypes.c_int
    if path is None:
        _library = lib
    return lib


class NativeReference:
    """An owned reference handle (visualizer.py:6250-6370, without the Python
    fallback tiers)."""

    def __init__(self, x_center: str, y_center: str, viewport_zoom: str, max_iter: int,
                 precision_bits: int, image_series_order: int = 8,
                 options: NativeRenderOptions | None = None, library=None, formula: int = 0,
                 julia: tuple[str, str] = ("0", "0")):
        self.library = library or load_library()
        self.options = options or NativeRenderOptions()
        self.formula = int(formula)
        handle = self.library.fractal_create_reference_ex_options(
            x_center.encode(), y_center.encode(), viewport_zoom.encode(), int(max_iter),
            int(precision_bits), int(image_series_order), self.formula, str(julia[0]).encode(),
            str(julia[1]).encode(), ctypes.byref(self.options))
        if not handle:
            raise RuntimeError(self.library.fractal_last_error().decode())
        self.handle = handle

    def upload(self) -> None:
        if self.library.fractal_hip_reference_upload(self.handle) != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())

    def render(self, width: int, height: int, zoom: str, max_iter: int, series_order: int = 3,
               series_block: int = 256, planes: bool = False,
               options: NativeRenderOptions | None = None):
        """Field (H, W) float32 and, when requested, a dict of planes."""
        opts = options or self.options
        n = int(width) * int(height)
        out = np.empty(n, dtype=np.float32)
        plane_arrays = None
        plane_struct = None
        if planes:
            plane_arrays = {"orbit_iteration": np.empty(n, dtype=np.int64)}
            for name in ("phase", "de_x", "de_y", "test1", "test2"):
                plane_arrays[name] = np.empty(n, dtype=np.float64)
            plane_struct = NativeRenderPlanes()
            plane_struct.struct_size = ctypes.sizeof(NativeRenderPlanes)
            plane_struct.version = 1
            plane_struct.orbit_iteration = plane_arrays["orbit_iteration"].ctypes.data_as(_c_int64_p)
            for name in ("phase", "de_x", "de_y", "test1", "test2"):
                setattr(plane_struct, name, plane_arrays[name].ctypes.data_as(_c_double_p))
        status = self.library.fractal_render_reference_ex_planes(
            out.ctypes.data_as(_c_float_p), int(width), int(height), zoom.encode(), self.handle,
            int(max_iter), 0, int(series_order), int(series_block), ctypes.byref(opts),
            ctypes.byref(plane_struct) if plane_struct is not None else None)
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        field = out.reshape(height, width)
        if planes:
            return field, {k: v.reshape(height, width) for k, v in plane_arrays.items()}
        return field

    def render_points(self, real_mantissa, imag_mantissa, exponents, zoom: str, max_iter: int,
                      series_order: int = 3, series_block: int = 256,
                      options: NativeRenderOptions | None = None) -> np.ndarray:
        """float32 value per point of the offsets dc = (real, imag) * 2^exponent
        from this reference's centre (fractal_render_points, renderer.cpp:
        20019-20186; the call visualizer.py:6674-6790 makes for glitch cells)."""
        re = np.ascontiguousarray(real_mantissa, dtype=np.float64).ravel()
        im = np.ascontiguousarray(imag_mantissa, dtype=np.float64).ravel()
        e = np.ascontiguousarray(exponents, dtype=np.int32).ravel()
        if not (re.size == im.size == e.size):
            raise ValueError("point mantissas and exponents must have one entry per point")
        out = np.empty(re.size, dtype=np.float32)
        opts = options or self.options
        status = self.library.fractal_render_points(
            out.ctypes.data_as(_c_float_p), int(re.size), zoom.encode(), re.ctypes.data_as(_c_double_p),
            im.ctypes.data_as(_c_double_p), e.ctypes.data_as(_c_int32_p), self.handle, int(max_iter), 0,
            int(series_order), int(series_block), ctypes.byref(opts))
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        return out

    def render_scheduled(self, width: int, height: int, zoom: str, max_iter: int, tile_order=None,
                         work: bool = False, series_order: int = 3, series_block: int = 256,
                         options: NativeRenderOptions | None = None):
        """Field (H, W) of any formula (zoom >= 1e6, no planes) with its
        32x32 tiles launched in `tile_order` (None = the plain grid); with
        work=True also the (H, W) int32 map of per-pixel loop passes."""
        opts = options or self.options
        out = np.empty((int(height), int(width)), dtype=np.float32)
        work_map = np.empty((int(height), int(width)), dtype=np.int32) if work else None
        order = None if tile_order is None else np.ascontiguousarray(tile_order, dtype=np.uint32)
        status = self.library.fractal_hip_render_reference_scheduled(
            out.ctypes.data_as(_c_float_p), int(width), int(height), zoom.encode(), self.handle, int(max_iter),
            int(series_order), int(series_block), ctypes.byref(opts),
            order.ctypes.data if order is not None else None, 0 if order is None else order.size,
            work_map.ctypes.data if work_map is not None else None)
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        return (out, work_map) if work else out

    def render_floatexp(self, width: int, height: int, zoom: str, max_iter: int, series_order: int = 3,
                        series_block: int = 256, options: NativeRenderOptions | None = None):
        """The FP32 floatexp field (H, W) and its per-pixel escape counts
        (max_iter inside, -1 unresolved), any formula."""
        return self._field_and_counts(self.library.fractal_hip_render_reference_floatexp, width, height, zoom,
                                      max_iter, series_order, series_block, options, 0)

    def render_counts(self, width: int, height: int, zoom: str, max_iter: int, series_order: int = 3,
                      series_block: int = 256, options: NativeRenderOptions | None = None):
        """The FP64 field (H, W) and its per-pixel escape counts, the
        convention of render_floatexp."""
        return self._field_and_counts(self.library.fractal_hip_render_reference_counts, width, height, zoom,
                                      max_iter, series_order, series_block, options)

    def _field_and_counts(self, entry, width, height, zoom, max_iter, series_order, series_block, options, *extra):
        opts = options or self.options
        out = np.empty((int(height), int(width)), dtype=np.float32)
        counts = np.empty((int(height), int(width)), dtype=np.int32)
        status = entry(out.ctypes.data, counts.ctypes.data, int(width), int(height), zoom.encode(), self.handle,
                       int(max_iter), int(series_order), int(series_block), ctypes.byref(opts), *extra)
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        return out, counts

    def preview_features(self, width: int, height: int, zoom: str, max_iter: int, work_limit: int,
                         series_order: int = 3, series_block: int = 256,
                         options: NativeRenderOptions | None = None) -> np.ndarray:
        """(tiles, 65) float32 cost-model inputs of a field from its
        quarter-size preview: at most `work_limit` loop passes per pixel for
        the alternate formulas, cap min(max_iter, work_limit) for the
        Mandelbrot map."""
        opts = options or self.options
        tiles = ((int(width) + 31) // 32) * ((int(height) + 31) // 32)
        out = np.empty((tiles, 65), dtype=np.float32)
        status = self.library.fractal_hip_preview_features(
            out.ctypes.data_as(_c_float_p), int(width), int(height), zoom.encode(), self.handle, int(max_iter),
            int(series_order), int(series_block), ctypes.byref(opts), int(work_limit))
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        return out

    def render_learned(self, width: int, height: int, zoom: str, max_iter: int, weights: np.ndarray,
                       work_limit: int, series_order: int = 3, series_block: int = 256,
                       options: NativeRenderOptions | None = None):
        """Field (H, W) of any formula rendered in the cost model's
        heaviest-first tile order; returns (field, order, timings dict)."""
        opts = options or self.options
        w = np.ascontiguousarray(weights, dtype=np.float32)
        out = np.empty((int(height), int(width)), dtype=np.float32)
        tiles = ((int(width) + 31) // 32) * ((int(height) + 31) // 32)
        order = np.empty(tiles, dtype=np.uint32)
        timings = (ctypes.c_double * 5)()
        status = self.library.fractal_hip_render_reference_learned(
            out.ctypes.data_as(_c_float_p), int(width), int(height), zoom.encode(), self.handle, int(max_iter),
            int(series_order), int(series_block), ctypes.byref(opts), w.ctypes.data_as(_c_float_p), int(work_limit),
            order.ctypes.data, timings)
        if status != 0:
            raise RuntimeError(self.library.fractal_last_error().decode())
        names = ("preview_ms", "model_ms", "render_ms", "wall_to_launch_ms", "wall_ms")
        return out, order, dict(zip(names, list(timings)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.library.fractal_destroy_reference(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class NativeBatchField(ctypes.Structure):
    """FractalHipBatchField (include/fractal_hip.h)."""

    _fields_ = [
        ("handle", ctypes.c_void_p), ("zoom_text", ctypes.c_char_p), ("device_output", ctypes.c_void_p),
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("max_iter", ctypes.c_int32),
        ("series_order", ctypes.c_int32), ("series_block", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class NativeSegmentField(ctypes.Structure):
    """FractalHipSegmentField (include/fractal_hip.h)."""

    _fields_ = [
        ("x_center", ctypes.c_char_p), ("y_center", ctypes.c_char_p), ("zoom_text", ctypes.c_char_p),
        ("device_output", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("max_iter", ctypes.c_int32), ("precision_bits", ctypes.c_int32), ("series_order", ctypes.c_int32),
        ("series_block", ctypes.c_int32),
    ]


def render_segment(fields, outputs, options: "NativeRenderOptions", threads: int = 16, pipelined: bool = True,
                   series_order: int = 3, series_block: int = 256, library=None) -> dict:
    """Builds each field's reference on `threads` host threads and renders
    the fields into the device buffers `outputs` (fields: dicts with x, y,
    zoom, width, height, max_iter, bits), pipelined or build-then-render.
    Returns the call's timings in ms."""
    lib = library or load_library()
    entries = (NativeSegmentField * len(fields))()
    keep = []
    for i, f in enumerate(fields):
        texts = (f["x"].encode(), f["y"].encode(), f["zoom"].encode())
        keep.append(texts)
        entries[i] = NativeSegmentField(texts[0], texts[1], texts[2], outputs[i], f["width"], f["height"],
                                        f["max_iter"], f["bits"], series_order, series_block)
    timings = (ctypes.c_double * 4)()
    status = lib.fractal_hip_render_segment(entries, len(fields), ctypes.byref(options), threads, int(pipelined),
                                            timings)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return dict(zip(("wall_ms", "build_ms", "render_ms", "built_ms"), list(timings)))


class HipRuntime:
    """Device memory through the HIP runtime that libfractal_hip.so links
    (loading a second HIP runtime into the process, e.g. another copy
    bundled with a framework, leaves that one without devices)."""

    def __init__(self):
        load_library()
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def malloc(self, nbytes: int) -> ctypes.c_void_p:
        ptr = ctypes.c_void_p()
        status = self.lib.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes))
        if status != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed with {status}")
        return ptr

    def free(self, ptr) -> None:
        self.lib.hipFree(ptr)

    def to_host(self, host_array: np.ndarray, ptr) -> None:
        status = self.lib.hipMemcpy(host_array.ctypes.data_as(ctypes.c_void_p), ptr, host_array.nbytes, 2)
        if status != 0:
            raise RuntimeError(f"hipMemcpy D2H failed with {status}")

    def to_device(self, ptr, host_array: np.ndarray) -> None:
        source = np.ascontiguousarray(host_array)
        status = self.lib.hipMemcpy(ptr, source.ctypes.data_as(ctypes.c_void_p), source.nbytes, 1)
        if status != 0:
            raise RuntimeError(f"hipMemcpy H2D failed with {status}")

    def synchronize(self) -> None:
        self.lib.hipDeviceSynchronize()


def colourise_kfp(field: np.ndarray, max_iter: int, options: NativeKfpOptions, lut: np.ndarray,
                  library=None) -> np.ndarray:
    """(H, W, 3) uint8 Kalles colouring of a float32 field on the GPU."""
    lib = library or load_library()
    height, width = field.shape
    source = np.ascontiguousarray(field, dtype=np.float32)
    table = np.ascontiguousarray(lut, dtype=np.uint8)
    out = np.empty((height, width, 3), dtype=np.uint8)
    status = lib.fractal_colourise_kfp(source.ctypes.data_as(_c_float_p), out.ctypes.data_as(_c_uint8_p), width,
                                       height, int(max_iter), 0.0, 0.0, 0.0, 0.0, ctypes.byref(options),
                                       table.ctypes.data_as(_c_uint8_p), table.shape[0], 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def last_timing() -> dict:
    """HIP-event kernel time and wall time of the most recent render, and the
    plain pass's own HIP-event time (0 when it did not run)."""
    values = (ctypes.c_double * 4)()
    load_library().fractal_hip_last_timing(values, 4)
    return {"kernel_ms": values[0], "total_ms": values[1], "pixels": values[2], "plain_ms": values[3]}
