"""The pipelined segment entry (fractal_hip_render_segment): references
built on host threads, tables staged in pinned memory and copied on a side
stream while earlier fields render.  Every field equals its single-field
handle render, pipelined or not, direct-renderer fields included."""
import numpy as np
import pytest

import fav_batch
import fav_hip

pytestmark = pytest.mark.gpu


def test_pipelined_segment_equals_single_field_renders(gpu_product):
    # zooms 10^(40k/11): k = 0, 1 are direct-renderer fields
    fields = fav_batch.segment_fields(count=12, seed=3, decades=40.0, sizes=((96, 54), (160, 90), (200, 120)),
                                      cap_scale=0.25)
    print(f"[gpu] pipelined segment of {len(fields)} fields", flush=True)
    batch = fav_batch.Batch(fields, threads=4)
    try:
        batch.render_per_field()
        singles = [batch.field(i).copy() for i in range(len(fields))]
        for pipelined in (True, False):
            for ptr, f in zip(batch.outputs, fields):
                batch.hip.to_device(ptr, np.full(f["width"] * f["height"], np.nan, dtype=np.float32))
            timings = fav_hip.render_segment(fields, batch.outputs, batch.options, threads=4, pipelined=pipelined)
            print(f"[gpu] pipelined {pipelined}: {timings}", flush=True)
            assert timings["wall_ms"] > 0 and timings["render_ms"] > 0 and timings["build_ms"] > 0
            for i, expected in enumerate(singles):
                np.testing.assert_array_equal(batch.field(i).view(np.uint32), expected.view(np.uint32),
                                              err_msg=f"pipelined {pipelined} field {i}")
    finally:
        batch.close()


def test_segment_entry_rejects_bad_fields(gpu_product):
    lib = fav_hip.load_library()
    opts = fav_hip.NativeRenderOptions(series_min_terms=32, escape_radius_mode=1, coordinate_mode=1)
    bad = [dict(x="-0.75", y="0.1", zoom="1e10", width=0, height=10, max_iter=100, bits=128)]
    with pytest.raises(RuntimeError, match="invalid"):
        fav_hip.render_segment(bad, [None], opts, library=lib)
