"""HIP direct renderer (use_perturbation = 0) against the reference's golden
vectors and the C oracle.

Bar: iteration plane, test1/test2 and interior pixels bit-exact; the float
field within one ulp (the reference subtracts in x87 long double); phase
and DE within 1e-9 relative."""
import numpy as np
import pytest

import golden_io
import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.direct_case_names())
def test_hip_direct_matches_reference_golden(gpu_product, name):
    print(f"[gpu] direct golden {name}", flush=True)
    g = golden_io.load_direct(name)
    field, planes = native.render_direct_abi(gpu_product, g["meta"])
    golden_io.assert_direct_parity(g, field, planes)


def test_hip_render_mandelbrot_entries_agree(gpu_product):
    print("[gpu] render_mandelbrot / _ex / render_fractal_ex", flush=True)
    case = dict(width=96, height=54, zoom="3", x="-0.75", y="0.1", max_iter=700, formula=0, planes=False,
                options={})
    a, _ = native.render_direct_abi(gpu_product, case)
    b = np.empty(96 * 54, dtype=np.float32)
    assert gpu_product.render_mandelbrot(b.ctypes.data_as(native.c_float_p), 96, 54, b"3", b"-0.75", b"0.1", 700,
                                         128, 0, 0) == 0, gpu_product.fractal_last_error()
    c = np.empty(96 * 54, dtype=np.float32)
    opts = native.direct_options(case, 2)
    assert gpu_product.render_mandelbrot_ex(c.ctypes.data_as(native.c_float_p), 96, 54, b"3", b"-0.75", b"0.1",
                                            700, 128, 0, 0, native.ctypes.byref(opts)) == 0
    np.testing.assert_array_equal(a.reshape(-1), b)
    np.testing.assert_array_equal(a.reshape(-1), c)


@pytest.mark.parametrize("formula", [0, 1, 2, 3])
def test_hip_direct_full_hd_matches_oracle(gpu_product, oracle, formula):
    print(f"[gpu] direct 1920x1080 formula {formula}", flush=True)
    case = dict(width=1920, height=1080, zoom="1.3", x="-0.6" if formula != 1 else "0", y="0.05",
                max_iter=2000, formula=formula, julia_re=-0.4, julia_im=0.6, planes=formula in (0, 1),
                options=dict(escape_radius_mode=formula % 2))
    field, planes = native.render_direct_abi(gpu_product, case)
    ofield, oplanes = native.render_direct_oracle(oracle, case)
    golden = {"field": ofield}
    if oplanes is not None:
        golden.update(oplanes.as_dict())
    golden_io.assert_direct_parity(golden, field, planes)
