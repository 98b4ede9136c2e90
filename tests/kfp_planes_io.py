"""Kalles plane inputs of the plane-aware KFP colour pass
(fractal_colourise_kfp_planes, renderer.h:483-501): a FractalKfpPlanes
mirror owning numpy storage, the case inputs regenerated from the field
golden vectors and seeds, and the calls through a library or the CPU build
of the kernel code."""
import ctypes

import numpy as np

import native
from golden_io import GOLDEN
from kfp_cases import KALLES_DEFAULT_STOPS, seeded_stops

I64 = ctypes.POINTER(ctypes.c_int64)
F64 = ctypes.POINTER(ctypes.c_double)


class KfpPlanesStruct(ctypes.Structure):
    """FractalKfpPlanes, renderer.h:135-153."""

    _fields_ = [("struct_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
                ("orbit_iteration", I64), ("bailout", F64), ("iteration", I64), ("transition", F64),
                ("phase", F64), ("de_x", F64), ("de_y", F64), ("test1", F64), ("test2", F64),
                ("texture_rgb", native.c_uint8_p), ("texture_width", ctypes.c_int32),
                ("texture_height", ctypes.c_int32), ("texture_stride", ctypes.c_int32)]


class KfpPlanes:
    """Owns the arrays behind a KfpPlanesStruct (absent planes stay NULL)."""

    INT = ("orbit_iteration", "iteration")
    DOUBLE = ("bailout", "transition", "phase", "de_x", "de_y", "test1", "test2")

    def __init__(self, arrays, texture=None):
        self.arrays = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        self.texture = texture
        self.struct = KfpPlanesStruct()
        self.struct.struct_size = ctypes.sizeof(KfpPlanesStruct)
        self.struct.version = 1
        for name in self.INT:
            if name in self.arrays:
                self.arrays[name] = self.arrays[name].astype(np.int64)
                setattr(self.struct, name, self.arrays[name].ctypes.data_as(I64))
        for name in self.DOUBLE:
            if name in self.arrays:
                self.arrays[name] = self.arrays[name].astype(np.float64)
                setattr(self.struct, name, self.arrays[name].ctypes.data_as(F64))
        if texture is not None:
            rgb, width, height, stride = texture
            self.texture_bytes = np.ascontiguousarray(rgb, dtype=np.uint8)
            self.struct.texture_rgb = self.texture_bytes.ctypes.data_as(native.c_uint8_p)
            self.struct.texture_width, self.struct.texture_height, self.struct.texture_stride = width, height, stride


def texture_image(width=37, height=23, stride_pad=5, seed=3):
    rng = np.random.default_rng(seed)
    stride = width * 3 + stride_pad
    return rng.integers(0, 256, size=stride * height).astype(np.uint8), width, height, stride


def inputs(case, oracle):
    """(field, planes, lut, options, max_iter) of a case."""
    with np.load(GOLDEN / f"{case['source']}.npz", allow_pickle=False) as g:
        field = g["field"].astype(np.float32)
        render = {k: g[k].reshape(field.shape) for k in ("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")}
    rng = np.random.default_rng(case["seed"])
    arrays = {k: render[k] for k in case["planes"] if k in render}
    if "iteration" in case["planes"]:  # stored iteration + transition = 1 - fraction
        arrays["iteration"] = render["orbit_iteration"] + 1
        frac = np.nan_to_num(np.mod(field.astype(np.float64), 1.0))
        arrays["transition"] = 1.0 - frac
    if "transition_glitches" in case:
        t = arrays.setdefault("transition", np.ones(field.shape))
        t.reshape(-1)[rng.choice(t.size, size=case["transition_glitches"], replace=False)] = -1.0
    if "bailout" in case["planes"]:
        b = np.full(field.shape, 10000.0)
        b.reshape(-1)[rng.choice(b.size, size=b.size // 5, replace=False)] = 2.0
        b.reshape(-1)[rng.choice(b.size, size=b.size // 20, replace=False)] = np.nan
        arrays["bailout"] = b
    texture = texture_image() if case.get("texture") else None
    stops = np.asarray(KALLES_DEFAULT_STOPS if case["stops"] == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    opts = native.KfpOptions.make(**case["options"])
    return field, KfpPlanes(arrays, texture), lut, opts, case["max_iter"]


def colourise(lib, field, planes, lut, opts, max_iter):
    h, w = field.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    status = lib.fractal_colourise_kfp_planes(
        field.ctypes.data_as(native.c_float_p), out.ctypes.data_as(native.c_uint8_p), w, h, max_iter, 0.0, 0.0, 0.0,
        0.0, ctypes.byref(opts), ctypes.byref(planes.struct), lut.ctypes.data_as(native.c_uint8_p), lut.shape[0], 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def colourise_host_check(lib, field, planes, lut, opts, max_iter):
    lib.fav_host_check_kfp_planes.argtypes = [
        native.c_float_p, native.c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(native.KfpOptions), ctypes.POINTER(KfpPlanesStruct), native.c_uint8_p, ctypes.c_int,
        ctypes.c_char_p, ctypes.c_int]
    lib.fav_host_check_kfp_planes.restype = ctypes.c_int
    h, w = field.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    err = ctypes.create_string_buffer(256)
    status = lib.fav_host_check_kfp_planes(
        field.ctypes.data_as(native.c_float_p), out.ctypes.data_as(native.c_uint8_p), w, h, max_iter,
        ctypes.byref(opts), ctypes.byref(planes.struct), lut.ctypes.data_as(native.c_uint8_p), lut.shape[0], err, 256)
    if status != 0:
        raise RuntimeError(err.value.decode())
    return out


ATLAS_SIGNATURE = None


def render_tile(source, planes=("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2"), centred=False,
                cap=None, extra=None):
    """(field, KfpPlanes) of a field golden vector used as an atlas tile."""
    with np.load(GOLDEN / f"{source}.npz", allow_pickle=False) as g:
        field = g["field"].astype(np.float32)
        arrays = {k: g[k].reshape(field.shape) for k in planes}
        cap = cap or json_max_iter(g)
    if centred:
        field = (field.astype(np.float64) - cap).astype(np.float32)
    arrays.update(extra or {})
    return np.ascontiguousarray(field), arrays, cap


def json_max_iter(g):
    import json
    return int(json.loads(str(g["meta"]))["max_iter"])


def atlas_inputs(case, oracle):
    """Parent / child tiles with their planes, LUT and options of an atlas case."""
    rng = np.random.default_rng(case["seed"])
    centred = bool(case.get("centred", 0))
    pf, pa, pcap = render_tile(case["parent"], case.get("parent_planes", FULL_PLANES), centred)
    parent = (pf, KfpPlanes(pa, texture_image() if case.get("parent_texture") else None), pcap)
    child = None
    if case.get("child"):
        extra = {}
        cf0, _, _ = render_tile(case["child"], (), False)
        if case.get("child_transition_glitches"):
            t = np.ones(cf0.shape)
            t.reshape(-1)[rng.choice(t.size, size=case["child_transition_glitches"], replace=False)] = -1.0
            extra["transition"] = t
        cf, ca, ccap = render_tile(case["child"], case.get("child_planes", FULL_PLANES), centred, extra=extra)
        child = (cf, KfpPlanes(ca, texture_image(seed=9) if case.get("child_texture") else None), ccap)
    stops = np.asarray(KALLES_DEFAULT_STOPS if case["stops"] == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    opts = native.KfpOptions.make(**case["options"])
    return parent, child, lut, opts


FULL_PLANES = ("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")


def atlas_colourise(lib, case, parent, child, lut, opts, entry="fractal_atlas_colourise_kfp_planes"):
    fn = getattr(lib, entry)
    fn.argtypes = [native.c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, native.c_float_p,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, native.c_uint8_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(native.KfpOptions), native.c_uint8_p, ctypes.c_int, ctypes.c_int]
    fn.restype = ctypes.c_int
    ow, oh = case["ow"], case["oh"]
    out = np.zeros((oh, ow, 3), dtype=np.uint8)
    pf, pp, pcap = parent
    if child is not None:
        cf, cp, ccap = child
        child_args = (cf.ctypes.data_as(native.c_float_p), cf.shape[1], cf.shape[0], ccap,
                      ctypes.cast(ctypes.byref(cp.struct), ctypes.c_void_p))
    else:
        child_args = (None, 0, 0, 0, None)
    status = fn(pf.ctypes.data_as(native.c_float_p), pf.shape[1], pf.shape[0], pcap,
                ctypes.cast(ctypes.byref(pp.struct), ctypes.c_void_p), *child_args,
                out.ctypes.data_as(native.c_uint8_p), ow, oh, case["parent_zoom"], case["fraction"], case["child_zoom"],
                case["max_iter"], int(case.get("centred", 0)), 0.0, 0.0, 0.0, 0.0, ctypes.byref(opts),
                lut.ctypes.data_as(native.c_uint8_p), lut.shape[0], 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def crop_inputs(case, oracle):
    with np.load(GOLDEN / f"{case['source']}.npz", allow_pickle=False) as g:
        field = g["field"].astype(np.float64)
    if case["centred"]:
        field = field - case["max_iter"]
    stops = np.asarray(KALLES_DEFAULT_STOPS if case["stops"] == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    return np.ascontiguousarray(field, dtype=np.float32), lut, native.KfpOptions.make(**case["options"])


def crop_colourise(lib, case, field, lut, opts, entry="fractal_crop_colourise_kfp"):
    fn = getattr(lib, entry)
    fn.argtypes = [native.c_float_p, ctypes.c_int, ctypes.c_int, native.c_uint8_p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(native.KfpOptions), native.c_uint8_p, ctypes.c_int, ctypes.c_int]
    fn.restype = ctypes.c_int
    out = np.zeros((case["oh"], case["ow"], 3), dtype=np.uint8)
    status = fn(field.ctypes.data_as(native.c_float_p), field.shape[1], field.shape[0],
                out.ctypes.data_as(native.c_uint8_p), case["ow"], case["oh"], case["zoom"], case["max_iter"], 0.0, 0.0,
                0.0, 0.0, ctypes.byref(opts), lut.ctypes.data_as(native.c_uint8_p), lut.shape[0], 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def rect_inputs(case, oracle):
    def field(name):
        with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as g:
            return np.ascontiguousarray(g["field"], dtype=np.float32)
    parent = field(case["parent"])
    child = field(case["child"]) if case["child"] else None
    stops = np.asarray(KALLES_DEFAULT_STOPS if case["stops"] == "kalles" else seeded_stops(), dtype=np.uint8)
    lut = native.palette_lut_oracle(oracle, stops, 1024)
    return parent, child, lut, native.KfpOptions.make(**case["options"])


def rect_colourise(lib, case, parent, child, lut, opts, entry="fractal_atlas_colourise_kfp"):
    fn = getattr(lib, entry)
    fn.argtypes = [native.c_float_p, ctypes.c_int, ctypes.c_int, native.c_float_p, ctypes.c_int, ctypes.c_int,
                   native.c_uint8_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(native.KfpOptions), native.c_uint8_p, ctypes.c_int, ctypes.c_int]
    fn.restype = ctypes.c_int
    h, w = parent.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    ch, cw = child.shape if child is not None else (0, 0)
    status = fn(parent.ctypes.data_as(native.c_float_p), w, h,
                child.ctypes.data_as(native.c_float_p) if child is not None else None, cw, ch,
                out.ctypes.data_as(native.c_uint8_p), w, h, case["max_iter"], case["left"], case["top"],
                case["feather"], 0.0, 0.0, 0.0, 0.0, ctypes.byref(opts), lut.ctypes.data_as(native.c_uint8_p),
                lut.shape[0], 0)
    if status != 0:
        raise RuntimeError(lib.fractal_last_error().decode())
    return out


def rect_host_check(lib, case, parent, child, lut, opts):
    fn = lib.fav_host_check_kfp_rect
    fn.argtypes = [native.c_float_p, native.c_float_p, ctypes.c_int, ctypes.c_int, native.c_uint8_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(native.KfpOptions), native.c_uint8_p, ctypes.c_int]
    fn.restype = ctypes.c_int
    h, w = parent.shape
    out = np.zeros((h, w, 3), dtype=np.uint8)
    ch, cw = child.shape if child is not None else (0, 0)
    assert fn(parent.ctypes.data_as(native.c_float_p), child.ctypes.data_as(native.c_float_p) if child is not None
              else None, cw, ch, out.ctypes.data_as(native.c_uint8_p), w, h, case["max_iter"], case["left"],
              case["top"], case["feather"], ctypes.byref(opts), lut.ctypes.data_as(native.c_uint8_p), lut.shape[0]) == 0
    return out
