"""The HIP library's C ABI without a GPU: it loads, exports every symbol
include/fractal_hip.h declares, validates arguments like the reference
(renderer.cpp:4729-4765, 19814-19860), builds references on the host, and
fails loudly instead of falling back when no device is present."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import native

HEADER = Path(__file__).resolve().parent.parent / "include" / "fractal_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_hot_path_abi():
    names = declared_functions()
    for required in ("fractal_abi_version", "fractal_last_error", "fractal_create_reference",
                     "fractal_destroy_reference", "render_mandelbrot_reference",
                     "fractal_render_reference_ex_planes", "render_fractal_ex"):
        assert required in names


def test_library_exports_every_declared_symbol(product):
    for name in declared_functions():
        assert hasattr(product, name), name


def test_identity(product):
    assert product.fractal_abi_version() == 10
    assert product.fractal_render_options_version() == 2
    opts = native.RenderOptions()
    product.fractal_render_options_default(ctypes.byref(opts))
    assert opts.struct_size == ctypes.sizeof(native.RenderOptions) == 72
    assert (opts.version, opts.strict, opts.series_min_terms, opts.series_max_terms) == (2, 1, 8, 32)
    assert (opts.max_bla_length, opts.max_linear_bla_length, opts.backend) == (64, 4096, 2)


def test_capabilities_never_claim_cpu_backends(product):
    caps = product.fractal_backend_capabilities()
    assert caps & 3 == 0
    assert caps in (0, 12)


X = b"-0.743643887037158704752191506114774"
Y = b"0.131825904205311970493132056385139"


def test_reference_builds_on_host_and_reports_stats(product):
    handle = product.fractal_create_reference(X, Y, b"1e12", 2000, 128, 8)
    assert handle, product.fractal_last_error()
    values = np.zeros(5, dtype=np.uint64)
    assert product.fractal_get_reference_stats(ctypes.c_void_p(handle), values.ctypes.data_as(native.c_uint64_p), 5) == 5
    assert values[4] == 8 and values[3] > 1  # series order and validated iteration
    product.fractal_destroy_reference(ctypes.c_void_p(handle))
    assert product.fractal_last_error() == b""
    product.fractal_destroy_reference(ctypes.c_void_p(handle))  # stale handle
    assert b"invalid or already-destroyed" in product.fractal_last_error()
    product.fractal_destroy_reference(None)  # NULL is a no-op
    assert product.fractal_last_error() == b""


@pytest.mark.parametrize("zoom, max_iter, bits, order", [
    ("1e12", 20000, 128, 8),     # the series stops inside the first 2048-iteration chunk
    ("1e12", 20000, 128, 24),
    ("1e16", 30000, 192, 8),     # ... in the second chunk
    ("1e25", 50000, 256, 8),     # ... in the sixth
    ("1e25", 50000, 256, 16),
])
def test_reference_stats_match_oracle_series(product, oracle, zoom, max_iter, bits, order):
    """The image series' stopping iteration and order (built in chunks with
    the probes on several threads) equal the sequential restatement's."""
    handle = product.fractal_create_reference(X, Y, zoom.encode(), max_iter, bits, order)
    values = np.zeros(5, dtype=np.uint64)
    product.fractal_get_reference_stats(ctypes.c_void_p(handle), values.ctypes.data_as(native.c_uint64_p), 5)
    ref = native.OracleReference(oracle, X.decode(), Y.decode(), zoom, max_iter, bits, order)
    info = ref.info()
    assert int(values[3]) == info["series_iteration"] and int(values[4]) == info["series_order"]
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


@pytest.mark.parametrize("kwargs,message", [
    (dict(backend=0), b"HIP backend"),
    (dict(backend=1), b"HIP backend"),
    (dict(backend=3), b"unknown native render backend"),
    (dict(strict=2), b"strict render flag must be 0 or 1"),
    (dict(escape_radius_mode=2), b"unknown escape-radius mode"),
    (dict(time_budget_ms=-1), b"render time budget cannot be negative"),
    (dict(output_bias=-1.0), b"output bias"),
])
def test_render_option_validation(product, kwargs, message):
    handle = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    out = np.zeros(16, dtype=np.float32)
    opts = native.RenderOptions.defaults(**{"backend": 2, **kwargs})
    status = product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e12",
                                                 ctypes.c_void_p(handle), 200, 0, 8, 32, ctypes.byref(opts))
    assert status == 1
    assert message in product.fractal_last_error()
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


def test_invalid_handles_and_dimensions(product):
    out = np.zeros(16, dtype=np.float32)
    opts = native.RenderOptions.defaults(backend=2)
    assert product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e12",
                                               ctypes.c_void_p(0x7777), 200, 0, 8, 32, ctypes.byref(opts)) == 1
    assert b"invalid or already-destroyed" in product.fractal_last_error()
    handle = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    assert product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 0, 4, b"1e12",
                                               ctypes.c_void_p(handle), 200, 0, 8, 32, ctypes.byref(opts)) == 1
    assert product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e12",
                                               ctypes.c_void_p(handle), 201, 0, 8, 32, ctypes.byref(opts)) == 1
    assert b"exceeds prepared reference" in product.fractal_last_error()
    kalles = native.RenderOptions.defaults(backend=2, escape_radius_mode=1)
    assert product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e12",
                                               ctypes.c_void_p(handle), 200, 0, 8, 32, ctypes.byref(kalles)) == 1
    assert b"escape-radius mode" in product.fractal_last_error()
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


def test_invalid_reference_arguments(product):
    assert not product.fractal_create_reference(X, Y, b"1e12", 0, 128, 8)
    assert not product.fractal_create_reference(X, Y, b"1e12", 100, 64, 8)
    assert not product.fractal_create_reference(b"not-a-number", Y, b"1e12", 100, 128, 8)
    assert b"invalid real coordinate" in product.fractal_last_error()
    assert not product.fractal_create_reference(X, Y, b"-5", 100, 128, 8)
    assert not product.fractal_create_reference_ex(X, Y, b"1e12", 100, 128, 8, 4, b"0", b"0")
    assert not product.fractal_create_reference_ex(X, Y, b"1e12", 100, 128, 8, -1, b"0", b"0")


def test_clone_requires_reusable_root(product):
    plain = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    assert not product.fractal_clone_reference(ctypes.c_void_p(plain), b"1e13")
    assert b"reusable tier root" in product.fractal_last_error()
    root = product.fractal_create_reference_reusable(X, Y, b"1e12", 200, 128, 8)
    clone = product.fractal_clone_reference(ctypes.c_void_p(root), b"1e13")
    assert clone, product.fractal_last_error()
    for h in (plain, root, clone):
        product.fractal_destroy_reference(ctypes.c_void_p(h))


def test_render_without_device_fails_loudly(product):
    if product.fractal_backend_capabilities() & 8:
        pytest.skip("a device is present")
    handle = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    out = np.full(16, 7.0, dtype=np.float32)
    opts = native.RenderOptions.defaults(backend=2)
    status = product.fractal_render_reference_ex(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e12",
                                                 ctypes.c_void_p(handle), 200, 0, 8, 32, ctypes.byref(opts))
    assert status == 1 and product.fractal_last_error()
    assert np.all(out == 7.0)  # nothing was computed on the host
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


def _kfp_call(product, options, field=None, lut_size=1024, max_iter=100, planes=None, phase=0.0):
    field = np.full((4, 6), 10.0, dtype=np.float32) if field is None else field
    lut = np.zeros((max(lut_size, 1), 3), dtype=np.uint8)
    out = np.zeros((field.shape[0], field.shape[1], 3), dtype=np.uint8)
    return product.fractal_colourise_kfp_planes(
        field.ctypes.data_as(native.c_float_p), out.ctypes.data_as(native.c_uint8_p), field.shape[1],
        field.shape[0], max_iter, phase, 0.0, 0.0, 0.0, ctypes.byref(options), planes,
        lut.ctypes.data_as(native.c_uint8_p), lut_size, 0)


@pytest.mark.parametrize("change", [
    dict(iter_div=0.0), dict(color_method=12), dict(differences=8), dict(smooth_method=3), dict(power=0.0),
    dict(slope_power=-1.0), dict(field_bias=-1.0), dict(bailout_radius_preset=4), dict(interior_color=(0, 0, 256)),
    dict(multi_colors=[(0.0, 0)]), dict(multi_colors=[(1.0, 3)]), dict(show_glitches=2), dict(version=2)])
def test_kfp_invalid_options_rejected_before_device(product, change):
    """valid_kfp_options, renderer.cpp:7396-7454."""
    opts = native.KfpOptions.make(**change)
    assert _kfp_call(product, opts) == 1
    assert b"invalid native KFP colour" in product.fractal_last_error()


def test_kfp_rejects_bad_lut_and_phase(product):
    opts = native.KfpOptions.make()
    assert _kfp_call(product, opts, lut_size=1) == 1
    assert _kfp_call(product, opts, phase=float("nan")) == 1


def test_kfp_rejects_field_bias_above_cap(product):
    opts = native.KfpOptions.make(field_bias=101.0)
    assert _kfp_call(product, opts, max_iter=100) == 1
    assert b"field bias exceeds" in product.fractal_last_error()


def test_kfp_malformed_planes_rejected(product):
    """A planes struct without the FractalKfpPlanes size and version is
    rejected before any device work (valid_kfp_planes, renderer.cpp:7467-7491)."""
    planes = ctypes.create_string_buffer(256)
    assert _kfp_call(product, native.KfpOptions.make(), planes=ctypes.cast(planes, ctypes.c_void_p)) == 1
    assert b"invalid native KFP colour dimensions or controls" in product.fractal_last_error()


def test_package_kfp_options_mirror_header_layout():
    """fav_hip.NativeKfpOptions and the test mirror share renderer.h's layout."""
    import fav_hip
    assert ctypes.sizeof(fav_hip.NativeKfpOptions) == ctypes.sizeof(native.KfpOptions)
    for (name, _), (other, _) in zip(fav_hip.NativeKfpOptions._fields_, native.KfpOptions._fields_):
        assert name == other
        assert getattr(fav_hip.NativeKfpOptions, name).offset == getattr(native.KfpOptions, name).offset
    a = fav_hip.NativeKfpOptions.kalles_default()
    b = native.KfpOptions.make(iter_div=0.01, color_method=7, slopes=1, differences=3)
    assert bytes(a) == bytes(b)


def test_package_loader_binds_colour_entries(product):
    import fav_hip
    lib = fav_hip.load_library()
    assert lib.fractal_hip_colourise_kfp_device.restype is ctypes.c_int


@pytest.mark.parametrize("zoom,x,err", [("0", "0", b"invalid Mandelbrot zoom"), ("1e-400", "0", b"outside the supported"),
                                        ("1", "1,5", b"invalid real coordinate")])
def test_direct_render_validation_before_device(product, zoom, x, err):
    out = np.zeros(4, dtype=np.float32)
    opts = native.RenderOptions.defaults(backend=2)
    status = product.render_fractal_ex(out.ctypes.data_as(native.c_float_p), 2, 2, zoom.encode(), x.encode(), b"0",
                                       100, 128, 0, 0, 0, 0.0, 0.0, ctypes.byref(opts))
    assert status == 1
    assert err in product.fractal_last_error()


def test_direct_planes_need_mandelbrot(product):
    """renderer.cpp:20399-20405: plane renders off the scalar backend are Mandelbrot-only."""
    planes = native.PlaneArrays(2, 2)
    out = np.zeros(4, dtype=np.float32)
    opts = native.RenderOptions.defaults(backend=2)
    status = product.render_fractal_ex_planes(out.ctypes.data_as(native.c_float_p), 2, 2, b"1", b"0", b"0", 100, 128,
                                              0, 0, 2, 0.0, 0.0, ctypes.byref(opts), ctypes.byref(planes.struct))
    assert status == 1
    assert b"scalar backend for this formula" in product.fractal_last_error()


def test_alternate_reference_creation_on_host(product):
    """Alternate handles build their MPFR orbit and real 2x2 BLA on the host
    (no device needed) and reject plane renders on the GPU backend."""
    opts = native.RenderOptions.defaults(backend=2)
    for formula, julia in ((1, (b"-0.8", b"0.156")), (2, (b"0", b"0")), (3, (b"0", b"0"))):
        handle = product.fractal_create_reference_ex_options(b"-0.99", b"0.05", b"1e7", 500, 128, 8, formula,
                                                             julia[0], julia[1], ctypes.byref(opts))
        assert handle, product.fractal_last_error()
        planes = native.PlaneArrays(4, 4)
        out = np.zeros(16, dtype=np.float32)
        status = product.fractal_render_reference_ex_planes(out.ctypes.data_as(native.c_float_p), 4, 4, b"1e7",
                                                            ctypes.c_void_p(handle), 500, 0, 8, 32,
                                                            ctypes.byref(opts), ctypes.byref(planes.struct))
        assert status == 1
        assert b"scalar backend for this formula" in product.fractal_last_error()
        clone = product.fractal_clone_reference(ctypes.c_void_p(handle), b"1e9")
        assert clone, product.fractal_last_error()
        product.fractal_destroy_reference(ctypes.c_void_p(clone))
        product.fractal_destroy_reference(ctypes.c_void_p(handle))


def test_alternate_rejects_bad_julia_constant(product):
    opts = native.RenderOptions.defaults(backend=2)
    handle = product.fractal_create_reference_ex_options(b"0", b"0", b"1e7", 500, 128, 8, 1, b"x", b"0",
                                                         ctypes.byref(opts))
    assert not handle


def test_atlas_segment_workload_rules():
    """Keyframe zooms 10^(100k/47), caps 2000 + 1000/decade clamped to
    [64, 100000], sizes from the three frame sizes, all 48 fields (k = 0, 1,
    2 below 1e6 are direct-renderer fields)."""
    import fav_batch
    fields = fav_batch.segment_fields()
    assert len(fields) == 48
    assert sum(1 for f in fields if f["zoom_log10"] < 6.0) == 3
    for f in fields:
        assert f["max_iter"] == min(100000, max(64, int(2000 + 1000 * f["zoom_log10"])))
        assert (f["width"], f["height"]) in fav_batch.SIZES
        assert float(f["zoom"]) == pytest.approx(10.0 ** f["zoom_log10"], rel=1e-9)
    assert fields[-1]["zoom_log10"] == 100.0


def test_batch_creation_validates_without_device(product):
    import fav_hip
    lib = fav_hip.load_library()
    entries = (fav_hip.NativeBatchField * 1)()
    entries[0].handle = None
    entries[0].zoom_text = b"1e10"
    entries[0].device_output = None
    assert not lib.fractal_hip_batch_create(entries, 1, ctypes.byref(fav_hip.NativeRenderOptions()))
    assert not lib.fractal_hip_batch_create(entries, 0, ctypes.byref(fav_hip.NativeRenderOptions()))
    assert lib.fractal_hip_batch_tile_count(ctypes.c_void_p(0x7777)) == -1


def _points_call(product, handle, count=4, options=None, planes=None, zoom=b"1e12", max_iter=200, real=None,
                 exponents=None):
    re = np.full(count, 0.25) if real is None else real
    im = np.zeros(count)
    e = np.full(count, -42, dtype=np.int32) if exponents is None else exponents
    out = np.full(max(count, 1), 7.0, dtype=np.float32)
    opts = options if options is not None else native.RenderOptions.defaults(backend=2)
    args = (out.ctypes.data_as(native.c_float_p), count, zoom, re.ctypes.data_as(native.c_double_p),
            im.ctypes.data_as(native.c_double_p), e.ctypes.data_as(native.c_int32_p), ctypes.c_void_p(handle),
            max_iter, 0, 8, 32, ctypes.byref(opts))
    if planes is not None:
        return product.fractal_render_points_ex_planes(*args, planes), out
    return product.fractal_render_points(*args), out


@pytest.mark.parametrize("change, message", [
    (dict(count=0), b"invalid native point-render arguments"),
    (dict(count=100_000_001), b"invalid native point-render arguments"),
    (dict(max_iter=0), b"invalid native point-render limits"),
    (dict(max_iter=201), b"point render iteration count exceeds prepared reference"),
    (dict(options=dict(backend=0)), b"HIP backend"),
    (dict(options=dict(backend=5)), b"unknown native render backend"),
    (dict(options=dict(coordinate_mode=1)), b"requires project coordinates"),
    (dict(zoom=b"zoom"), b"zoom"),
    (dict(real=np.array([0.25, np.inf, 0.0, 0.0])), b"non-finite point offset"),
])
def test_point_render_validation(product, change, message):
    """fractal_render_points_impl's checks and messages (renderer.cpp:20034-20080),
    then the HIP point path's own limits; no device work is started."""
    handle = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    kwargs = dict(change)
    if "options" in kwargs:
        kwargs["options"] = native.RenderOptions.defaults(**{"backend": 2, **kwargs["options"]})
    count = kwargs.pop("count", 4)
    if count > 1000:  # the count check comes before any pointer is read
        status = product.fractal_render_points(None, count, b"1e12", None, None, None, ctypes.c_void_p(handle),
                                               200, 0, 8, 32, None)
        out = np.full(4, 7.0, dtype=np.float32)
    else:
        status, out = _points_call(product, handle, count=count, **kwargs)
    assert status == 1
    assert message in product.fractal_last_error()
    assert np.all(out == 7.0)
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


def test_point_render_rejects_planes_and_stale_handles(product):
    handle = product.fractal_create_reference(X, Y, b"1e12", 200, 128, 8)
    planes = native.PlaneArrays(4, 1)
    status, _ = _points_call(product, handle, planes=ctypes.byref(planes.struct))
    assert status == 1 and b"point metadata requires the native scalar backend" in product.fractal_last_error()
    status, _ = _points_call(product, 0x7777)
    assert status == 1 and b"invalid or already-destroyed" in product.fractal_last_error()
    product.fractal_destroy_reference(ctypes.c_void_p(handle))


@pytest.mark.parametrize("formula", [0, 1, 2, 3])
def test_point_render_without_device_fails_loudly(product, formula):
    if product.fractal_backend_capabilities() & 8:
        pytest.skip("a device is present")
    handle = product.fractal_create_reference_ex(X, Y, b"1e12", 200, 128, 8, formula, b"-0.8", b"0.156")
    assert handle, product.fractal_last_error()
    status, out = _points_call(product, handle)
    assert status == 1 and product.fractal_last_error()
    assert np.all(out == 7.0)
    product.fractal_destroy_reference(ctypes.c_void_p(handle))
