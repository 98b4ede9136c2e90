"""The field kernels built with a five-block derivative log
(_build/libfractal_hip_log5.so, test-only): the log fills after five
committed linear blocks, so the full-log replay and the owed-derivative
settle run on the GPU's 64-lane waves -- with the production 4096-block
log they need more blocks in one loop than any test field commits.  Every
golden case and the 640x360 zoom-1e12 frame must come out exactly as the
reference computes them.

Bar: as the production library's parity tests (escape counts, NaN mask,
test1/test2 planes and all RenderStats counters bit-exact; float field
within one ulp; phase/DE within 1e-9 relative)."""
import ctypes

import pytest

import golden_io
import native

pytestmark = pytest.mark.gpu

LOG5_SO = native.ROOT / "fractal-audio-viz_amd" / "_build" / "libfractal_hip_log5.so"


@pytest.fixture(scope="module")
def log5_product(gpu_product):
    if not LOG5_SO.exists():
        pytest.fail("_build/libfractal_hip_log5.so is not built: run __graft_entry__.build()")
    lib = native.bind_kfp(native._bind_render_abi(ctypes.CDLL(str(LOG5_SO))))
    assert lib.fractal_backend_capabilities() & 8, lib.fractal_last_error()
    return lib


@pytest.mark.parametrize("name", golden_io.case_names())
def test_log5_kernels_match_golden(log5_product, name):
    print(f"[gpu] log5 golden case {name}", flush=True)
    golden = golden_io.load(name)
    field, planes, stats = golden_io.render_abi(log5_product, golden["meta"])
    golden_io.assert_hip_parity(golden, field, planes, stats)


def test_log5_full_640x360_matches_oracle(log5_product, oracle):
    print("[gpu] log5 640x360 1e12 full frame", flush=True)
    meta = dict(golden_io.load("c1_series")["meta"], width=640, height=360)
    f1, p1, s1 = golden_io.render_abi(log5_product, meta)
    f2, p2, s2 = golden_io.render_oracle(oracle, meta)
    golden = {"field": f2, "orbit_iteration": p2.orbit_iteration, "stats": s2,
              **{k: getattr(p2, k) for k in ("phase", "de_x", "de_y", "test1", "test2")}}
    golden_io.assert_hip_parity(golden, f1, p1, s1)
