"""The field kernel's per-pixel code (field_core.h), compiled for the CPU by
the test-only harness, against the golden vectors.  This checks the
kernel's control flow and table layouts without a GPU; the device build
is checked by test_gpu_parity.py."""
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
