"""The ordinary-palette colouriser and the per-frame crop / atlas
compositors on the CPU: the C oracle (oracle/fav_oracle_aurora.c) against
the fixtures the reference's renderer.cpp wrote, and the HIP library's
argument checks and its refusal to compute without a device."""
import ctypes

import numpy as np
import pytest

import aurora_io
from aurora_cases import ACCENTS, CASES


class OracleAurora:
    """The oracle's fo_X under the renderer.h name fractal_X."""

    def __init__(self, lib):
        self.lib = lib

    def __getattr__(self, name):
        fn = getattr(self.lib, "fo_" + name[len("fractal_"):])
        fn.argtypes = aurora_io.SIGNATURES[name]
        fn.restype = ctypes.c_int
        return fn


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fixture_inputs_regenerate(case):
    assert aurora_io.checksum(aurora_io.inputs(case)) == aurora_io.load(case["name"])["meta"]["inputs_sha256"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference(oracle, case):
    status, out = aurora_io.run(OracleAurora(oracle), case)
    assert status == 0
    expected = aurora_io.load(case["name"])["output"]
    if out.dtype == np.float32:
        np.testing.assert_array_equal(out.view(np.uint32), expected.view(np.uint32))
    else:
        np.testing.assert_array_equal(out, expected)


def test_fixtures_cover_the_branches():
    """Interior fills, every seam mode and the palette range are exercised."""
    full = aurora_io.load("aurora_atlas_colourise_interior_full")["output"]
    assert (full == np.array([9, 99, 199], dtype=np.uint8)).all(axis=-1).any()
    field = aurora_io.load("aurora_atlas_field_ex_centred")["output"]
    assert np.isfinite(field).all() and field.min() < -1000.0  # stored relative to the 4000 cap
    colours = aurora_io.load("aurora_colourise")["output"].reshape(-1, 3)
    assert len(np.unique(colours, axis=0)) > 100


def test_palette_pitch_half_is_unrotated(oracle):
    """At pitch 0.5 the hue angle is zero: the accents pass is the plain
    projection through the accent table (renderer.cpp:4813-4816)."""
    rgb = aurora_io.rgb(4, 8, 6)
    out = rgb.copy()
    acc = np.asarray(ACCENTS, dtype=np.uint8)
    OracleAurora(oracle).fractal_apply_aurora_accents(out.ctypes.data_as(aurora_io.U8), 8, 6,
                                                      acc.ctypes.data_as(aurora_io.U8), 0.5, 0)
    w = np.clip(rgb.astype(np.float64) / 140.0, 0.0, 1.0)
    expected = (w[..., 0:1] * acc[0:3] + w[..., 1:2] * acc[3:6] + w[..., 2:3] * acc[6:9]) / 1.8
    np.testing.assert_array_equal(out, np.rint(np.clip(expected, 0, 255)).astype(np.uint8))


@pytest.fixture(scope="module")
def hip(product):
    return aurora_io.bind(product)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_hip_without_device_fails_loudly(hip, case):
    if hip.fractal_backend_capabilities() & 8:
        pytest.skip("a device is present")
    status, out = aurora_io.run(hip, case)
    assert status == 1 and hip.fractal_last_error()
    if case["kind"] != "accents":
        assert not out.any()


@pytest.mark.parametrize("fn, args, message", [
    ("fractal_colourise", dict(max_iter=0), b"invalid native colour dimensions or palette"),
    ("fractal_colourise", dict(pitch=float("nan")), b"invalid native colour dimensions or palette"),
    ("fractal_colourise_accents", dict(accents=None), b"ordinary palette accents are required"),
    ("fractal_crop_rgb", dict(zoom=0.0), b"invalid native RGB crop dimensions"),
    ("fractal_atlas_composite_rgb", dict(fraction=1.5), b"invalid native RGB atlas dimensions or controls"),
    ("fractal_atlas_composite_rgb", dict(child=False, fraction=0.3), b"child data is missing"),
    ("fractal_crop_field", dict(sw=0), b"invalid native raw-field crop dimensions"),
    ("fractal_atlas_field_ex", dict(pbias=5000.0), b"invalid native scalar atlas dimensions"),
    ("fractal_atlas_field_ex", dict(obias=9000.0), b"invalid native scalar atlas iteration cap"),
    ("fractal_crop_colourise_interior", dict(interior=(0, 256, 0)), b"invalid native crop/colour dimensions"),
    ("fractal_atlas_colourise", dict(palette_max_iter=-1), b"invalid native atlas dimensions or controls"),
    ("fractal_atlas_colourise_accents", dict(accents=None), b"ordinary palette accents are required"),
])
def test_hip_argument_checks(hip, fn, args, message):
    case = dict(next(c for c in CASES if c["fn"] == fn))
    if "accents" in args and args["accents"] is None:
        case.pop("accents", None)
        args = {}
        case["fn"] = fn
    case.update(args)
    status, _ = aurora_io.run(hip, case)
    assert status == 1
    assert message in hip.fractal_last_error()
