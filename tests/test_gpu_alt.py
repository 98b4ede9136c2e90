"""HIP alternate deep formulas (Julia, burning ship, tricorn) against the
reference's golden vectors and the C oracle; shallow handle renders against
the direct renderer.

Bar: NaN mask, interior set and integer escape index bit-exact; float
field within one ulp (the reference's last logarithm is x87 long double)."""
import ctypes

import numpy as np
import pytest

import golden_io
import native
from alt_cases import BS_X, BS_Y, TC_X, TC_Y

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.alt_case_names())
def test_hip_alt_matches_reference_golden(gpu_product, name):
    print(f"[gpu] alternate golden {name}", flush=True)
    g = golden_io.load(name)
    field, stats = golden_io.render_alt_abi(gpu_product, g["meta"])
    golden_io.assert_alt_parity(g, field)
    assert int(stats[0]) == g["meta"]["width"] * g["meta"]["height"]


@pytest.mark.parametrize("formula,x,y", [(2, BS_X, BS_Y), (3, TC_X, TC_Y)])
def test_hip_alt_640x360_matches_oracle(gpu_product, oracle, formula, x, y):
    print(f"[gpu] alternate 640x360 formula {formula}", flush=True)
    case = dict(name="full", formula=formula, x=x, y=y, ref_zoom="4e7", zoom="4e7", max_iter=5000, bits=200,
                width=640, height=360, julia=["0", "0"], clone_zoom=None, options={})
    field, _ = golden_io.render_alt_abi(gpu_product, case)
    ofield, _ = golden_io.render_alt_oracle(oracle, case)
    golden_io.assert_alt_parity({"field": ofield, "meta": case}, field)


@pytest.mark.parametrize("formula", [0, 1, 2, 3])
def test_hip_shallow_handle_render_is_the_direct_renderer(gpu_product, formula):
    """renderer.cpp:19920-19960: a handle rendered below zoom 1e6 runs the
    direct double renderer at the handle's centre."""
    print(f"[gpu] shallow handle formula {formula}", flush=True)
    x, y = ("-0.7453", "0.1127") if formula != 1 else ("0.1", "0.2")
    julia = ("-0.8", "0.156") if formula == 1 else ("0", "0")
    opts = native.RenderOptions.defaults(backend=2)
    ref = native.AbiReference(gpu_product, x, y, "1e3", 1500, 128, 8, opts, formula=formula, julia=julia)
    a, _, _ = ref.render(96, 54, "1e3", 1500, 8, 32, opts)
    case = dict(width=96, height=54, zoom="1e3", x=x, y=y, max_iter=1500, formula=formula,
                julia_re=float(julia[0]), julia_im=float(julia[1]), planes=False, options={})
    b, _ = native.render_direct_abi(gpu_product, case)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
