"""Loading and comparing golden fixtures (tests/golden/*.npz)."""
import ctypes
import json
from pathlib import Path

import numpy as np

import native
from cases import CASES

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        out = {k: data[k] for k in data.files}
    out["meta"] = json.loads(str(out["meta"]))
    return out


def case_names():
    return [c["name"] for c in CASES]


def render_oracle(lib, case, stats=True):
    """Replays a golden case through the C oracle."""
    m = case
    ref = native.OracleReference(lib, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"],
                                 m["image_series_order"], m["options"].get("escape_radius_mode", 0),
                                 m["options"].get("coordinate_mode", 0))
    if m["clone_zoom"]:
        ref = ref.clone(m["clone_zoom"])
    opts = native.RenderOptions.defaults(**m["options"])
    return ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"],
                      m["series_block"], opts, planes=True, stats=stats)


def render_abi(lib, case, stats=True, backend=2):
    """Replays a golden case through a renderer.h-speaking library."""
    m = case
    opts = native.RenderOptions.defaults(backend=backend, **m["options"])
    if m["clone_zoom"]:
        handle = lib.fractal_create_reference_reusable_options(
            m["x"].encode(), m["y"].encode(), m["ref_zoom"].encode(), m["max_iter"], m["bits"],
            m["image_series_order"], ctypes.byref(opts))
        assert handle, lib.fractal_last_error()
        clone = lib.fractal_clone_reference(ctypes.c_void_p(handle), m["clone_zoom"].encode())
        assert clone, lib.fractal_last_error()
        lib.fractal_destroy_reference(ctypes.c_void_p(handle))
        ref = native.AbiReference.__new__(native.AbiReference)
        ref.lib, ref.handle = lib, clone
    else:
        ref = native.AbiReference(lib, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"],
                                  m["image_series_order"], opts)
    return ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"],
                      m["series_block"], opts, planes=True, stats=stats)


def stats_without_time(stats):
    s = np.asarray(stats, dtype=np.uint64).copy()
    s[22] = 0  # render_ns is wall time
    return s


def assert_bit_exact(golden, field, planes, stats):
    """Field, planes and counters identical to the last bit."""
    np.testing.assert_array_equal(field.reshape(-1).view(np.uint32), golden["field"].reshape(-1).view(np.uint32))
    np.testing.assert_array_equal(planes.orbit_iteration, golden["orbit_iteration"])
    for name in ("phase", "de_x", "de_y", "test1", "test2"):
        np.testing.assert_array_equal(getattr(planes, name).view(np.uint64), golden[name].view(np.uint64),
                                      err_msg=name)
    np.testing.assert_array_equal(stats_without_time(stats), stats_without_time(golden["stats"]))


# Tolerances of the HIP path against the FP64 CPU path.  Escape counts, the
# NaN mask, test1/test2 and all work counters are bit-exact; the float32
# field is bit-exact except that the reference forms the final smoothing
# logarithm in x87 long double (renderer.cpp:8170-8181), so one float32 ulp
# is allowed there; phase/DE use libm transcendentals and long double sums
# (renderer.cpp:7845-7896), held to 1e-9 relative.
FIELD_ULPS = 1
PLANE_RTOL = 1e-9


def assert_hip_parity(golden, field, planes, stats):
    f = field.reshape(-1)
    g = golden["field"].reshape(-1)
    np.testing.assert_array_equal(np.isnan(f), np.isnan(g))
    finite = ~np.isnan(g)
    ulps = np.abs(f[finite].view(np.int32).astype(np.int64) - g[finite].view(np.int32).astype(np.int64))
    assert ulps.size == 0 or int(ulps.max()) <= FIELD_ULPS, f"field differs by {int(ulps.max())} ulp"
    np.testing.assert_array_equal(planes.orbit_iteration, golden["orbit_iteration"])
    for name in ("test1", "test2"):
        np.testing.assert_array_equal(getattr(planes, name).view(np.uint64), golden[name].view(np.uint64),
                                      err_msg=name)
    for name in ("phase", "de_x", "de_y"):
        np.testing.assert_allclose(getattr(planes, name), golden[name], rtol=PLANE_RTOL, atol=0.0,
                                   err_msg=name)
    np.testing.assert_array_equal(stats_without_time(stats), stats_without_time(golden["stats"]))


def load_kfp(name):
    """A KFP colour golden case: field, lut, FractalKfpOptions, max_iter, rgb."""
    with np.load(GOLDEN / f"kfp_{name}.npz", allow_pickle=False) as data:
        out = {k: data[k] for k in data.files}
    out["options"] = native.KfpOptions.from_dict(json.loads(str(out["options"])))
    out["max_iter"] = int(out["max_iter"])
    return out


def kfp_case_names():
    from kfp_cases import CASES as KFP_CASES
    return [c[0] for c in KFP_CASES]


def assert_rgb_within_one_level(rgb, expected, name=""):
    """Colour bar: every channel within one 8-bit level of the reference."""
    diff = np.abs(rgb.astype(np.int16) - expected.astype(np.int16))
    assert int(diff.max()) <= 1, f"{name}: max RGB difference {int(diff.max())} at {np.argwhere(diff > 1)[:5]}"


def load_direct(name):
    """A direct-render golden case: meta (the case dict), field, planes."""
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        out = {k: data[k] for k in data.files}
    out["meta"] = json.loads(str(out["meta"]))
    return out


def direct_case_names():
    from direct_cases import CASES as DIRECT_CASES
    return [c["name"] for c in DIRECT_CASES]


def assert_direct_parity(golden, field, planes, exact_field=False):
    """Direct-render bar: iteration plane, test1/test2 and the interior set
    bit-exact; the float field within one ulp (bit-exact when exact_field);
    phase and DE within 1e-9 relative."""
    f = field.reshape(-1)
    g = golden["field"].reshape(-1)
    if exact_field:
        np.testing.assert_array_equal(f.view(np.uint32), g.view(np.uint32))
    else:
        ulps = np.abs(f.view(np.int32).astype(np.int64) - g.view(np.int32).astype(np.int64))
        assert int(ulps.max()) <= 1, f"field differs by {int(ulps.max())} ulp"
        assert np.array_equal(np.floor(f), np.floor(g)) or int(ulps.max()) <= 1
    if "orbit_iteration" in golden:
        np.testing.assert_array_equal(planes.orbit_iteration, golden["orbit_iteration"])
        np.testing.assert_array_equal(planes.test1, golden["test1"])
        np.testing.assert_array_equal(planes.test2, golden["test2"])
        for name in ("phase", "de_x", "de_y"):
            np.testing.assert_allclose(getattr(planes, name), golden[name], rtol=1e-9, atol=1e-300, err_msg=name)


def alt_case_names():
    from alt_cases import CASES as ALT_CASES
    return [c["name"] for c in ALT_CASES]


def render_alt_abi(lib, case, backend=2, stats=True):
    """An alternate-formula case through a renderer.h-speaking library."""
    m = case
    opts = native.RenderOptions.defaults(backend=backend, **m["options"])
    ref = native.AbiReference(lib, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], 8, opts,
                              formula=m["formula"], julia=tuple(m["julia"]))
    if m["clone_zoom"]:
        clone = lib.fractal_clone_reference(ctypes.c_void_p(ref.handle), m["clone_zoom"].encode())
        assert clone, lib.fractal_last_error()
        ref = native.AbiReference.__new__(native.AbiReference)
        ref.lib, ref.handle = lib, clone
    field, _, st = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], 8, 32, opts, stats=stats)
    return field, st


def render_alt_oracle(lib, case):
    m = case
    o = m["options"]
    ref = native.OracleReference(lib, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], 8,
                                 o.get("escape_radius_mode", 0), o.get("coordinate_mode", 0),
                                 formula=m["formula"], julia=tuple(m["julia"]))
    if m["clone_zoom"]:
        ref = ref.clone(m["clone_zoom"])
    opts = native.RenderOptions.defaults(**o)
    field, _, st = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], 8, 32, opts, stats=True)
    return field, st


def assert_alt_parity(golden, field):
    """Alternate-formula bar: NaN mask and interior set bit-exact, integer
    escape index bit-exact, float field within one ulp."""
    f = np.asarray(field, dtype=np.float32).reshape(-1)
    g = golden["field"].reshape(-1)
    np.testing.assert_array_equal(np.isnan(f), np.isnan(g))
    ok = ~np.isnan(g)
    ulps = np.abs(f[ok].view(np.int32).astype(np.int64) - g[ok].view(np.int32).astype(np.int64))
    assert ulps.size == 0 or int(ulps.max()) <= 1, f"field differs by {int(ulps.max())} ulp"
    bias = golden["meta"]["options"].get("output_bias", 0.0)
    np.testing.assert_array_equal(np.floor(f[ok].astype(np.float64) + bias), np.floor(g[ok].astype(np.float64) + bias))
