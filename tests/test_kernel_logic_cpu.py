"""The field kernel's per-pixel code (field_core.h), compiled for the CPU by
the test-only harness, against the golden vectors.  This checks the
kernel's control flow and table layouts without a GPU; the device build
is checked by test_gpu_parity.py."""
import numpy as np
import pytest

import golden_io
import host_check


@pytest.fixture(scope="module")
def harness():
    if not host_check.HOST_CHECK_SO.exists():
        pytest.fail("libfav_host_check.so is not built: run __graft_entry__.build()")
    return host_check.load()


@pytest.mark.parametrize("name", golden_io.case_names())
def test_kernel_code_on_cpu_matches_golden(harness, name):
    golden = golden_io.load(name)
    field, planes, stats = host_check.render(harness, golden["meta"])
    golden_io.assert_hip_parity(golden, field, planes, stats)


@pytest.fixture(scope="module")
def log5_harness():
    if not host_check.HOST_CHECK_LOG5_SO.exists():
        pytest.fail("libfav_host_check_log5.so is not built: run __graft_entry__.build()")
    return host_check.load(host_check.HOST_CHECK_LOG5_SO)


@pytest.mark.parametrize("name", golden_io.case_names())
def test_deferred_derivative_replays_on_full_log(harness, log5_harness, name):
    """The block loop defers each block's derivative update to a per-wave
    log; with a five-block log the full-log replays run many times per
    pixel.  The planes match the goldens, and the distance-estimate planes
    are bit-identical to those of the 4096-block log."""
    golden = golden_io.load(name)
    field, planes, stats = host_check.render(log5_harness, golden["meta"])
    golden_io.assert_hip_parity(golden, field, planes, stats)
    _, wide, _ = host_check.render(harness, golden["meta"])
    for plane in ("de_x", "de_y", "phase"):
        np.testing.assert_array_equal(getattr(planes, plane).view(np.uint64), getattr(wide, plane).view(np.uint64),
                                      err_msg=plane)


@pytest.fixture(scope="module")
def kfp_harness(harness):
    return host_check.bind_kfp(harness)


@pytest.mark.parametrize("name", golden_io.kfp_case_names())
def test_kfp_kernel_code_on_cpu_matches_golden(kfp_harness, name):
    """kfp_core.h on the CPU: bit-exact with the reference's RGB (same libm)."""
    g = golden_io.load_kfp(name)
    rgb = host_check.colourise(kfp_harness, g["field"], g["max_iter"], g["options"], g["lut"])
    np.testing.assert_array_equal(rgb, g["rgb"])


@pytest.fixture(scope="module")
def direct_harness(harness):
    return host_check.bind_direct(harness)


@pytest.mark.parametrize("name", golden_io.direct_case_names())
def test_direct_kernel_code_on_cpu_matches_golden(direct_harness, name):
    """direct_core.h on the CPU: field, counts and test planes bit-exact."""
    g = golden_io.load_direct(name)
    field, planes = host_check.render_direct(direct_harness, g["meta"])
    golden_io.assert_direct_parity(g, field, planes, exact_field=True)


@pytest.fixture(scope="module")
def alt_harness(harness):
    return host_check.bind_alt(harness)


@pytest.mark.parametrize("name", golden_io.alt_case_names())
def test_alt_kernel_code_on_cpu_matches_golden(alt_harness, name):
    """alt_core.h and the product's alternate tables on the CPU: escape
    counts bit-exact, float field within one ulp (double vs long double
    final logarithm)."""
    g = golden_io.load(name)
    field = host_check.render_alt(alt_harness, g["meta"])
    golden_io.assert_alt_parity(g, field)


def test_normal_operand_fast_ops_equal_general_ops(harness):
    """sc_mul_nn / sc_add_nn / sc_norm_nn are bit-identical to the general
    ScaledComplex operations on normal operands (2 million random pairs with
    zero parts, cancellations and exponent gaps around the 60-binade cut)."""
    import ctypes
    harness.fav_host_check_nn_ops.argtypes = [ctypes.c_longlong, ctypes.c_ulonglong,
                                              ctypes.POINTER(ctypes.c_longlong)]
    out = (ctypes.c_longlong * 4)()
    harness.fav_host_check_nn_ops(2_000_000, 20261203, out)
    assert list(out) == [0, 0, 0, 0]
