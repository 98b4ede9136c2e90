"""The schedule hooks of the Mandelbrot field kernels on the GPU: the
per-pixel work map (loop passes) sums to the RenderStats loop passes, a
32x32 tile order never changes the field or the work map, the preview
features are the patch of the capped quarter-size preview, the learned
render is the plain field, and a batch in the cost model's order keeps
every field.

Bars: fields and work maps bit-exact; features within 1e-6 (float logs)."""
import ctypes

import numpy as np
import pytest

import fav_batch
import fav_hip
import fav_sched
import fav_workloads

pytestmark = pytest.mark.gpu

C1 = dict(fav_workloads.field_1e12_640x360(), formula=0)


def c1_reference(lib):
    return fav_sched.reference_for(C1, lib)


def render_with_stats(lib, ref, **kw):
    case = C1
    lib.fractal_set_stats_enabled(1)
    try:
        field, work = ref.render_scheduled(case["width"], case["height"], case["zoom"], case["max_iter"], work=True,
                                           series_order=8, series_block=32, options=fav_sched.case_options(case),
                                           **kw)
        stats = np.zeros(39, dtype=np.uint64)
        lib.fractal_get_last_stats_ex(stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 39)
    finally:
        lib.fractal_set_stats_enabled(0)
    return field, work, stats


def test_mandelbrot_work_map_is_the_loop_passes(gpu_product):
    print("[gpu] mandelbrot work map 640x360 zoom 1e12", flush=True)
    lib = fav_hip.load_library()
    ref = c1_reference(lib)
    w, h = C1["width"], C1["height"]
    plain = ref.render(w, h, C1["zoom"], C1["max_iter"], 8, 32, options=fav_sched.case_options(C1)).reshape(h, w)
    field, work, stats = render_with_stats(lib, ref)
    np.testing.assert_array_equal(field.view(np.uint32), plain.view(np.uint32))
    # BLA blocks (linear, cubic), exact steps (tail included), replayed steps
    assert int(work.sum()) == int(stats[3]) + int(stats[4]) + int(stats[5]) + int(stats[6])
    assert work.min() >= 1 and work.max() <= C1["max_iter"]
    ref.close()


def test_mandelbrot_tile_order_never_changes_field_or_work(gpu_product):
    print("[gpu] mandelbrot tile orders 640x360", flush=True)
    lib = fav_hip.load_library()
    ref = c1_reference(lib)
    w, h = C1["width"], C1["height"]
    field, work, _ = render_with_stats(lib, ref)
    tiles = int(np.prod(fav_sched.tile_grid(w, h)))
    for order in (np.arange(tiles)[::-1], np.random.default_rng(5).permutation(tiles)):
        f2, w2, _ = render_with_stats(lib, ref, tile_order=order)
        np.testing.assert_array_equal(f2.view(np.uint32), field.view(np.uint32))
        np.testing.assert_array_equal(w2, work)
    ref.close()


@pytest.mark.parametrize("cap", [300, 4096])
def test_mandelbrot_preview_features_are_the_patch_of_the_capped_preview(gpu_product, cap):
    print(f"[gpu] mandelbrot preview features cap {cap}", flush=True)
    lib = fav_hip.load_library()
    ref = c1_reference(lib)
    w, h, z, n = C1["width"], C1["height"], C1["zoom"], C1["max_iter"]
    opts = fav_sched.case_options(C1)
    pw, ph = fav_sched.preview_size(w, h)
    _, preview_work = ref.render_scheduled(pw, ph, z, min(n, cap), work=True, series_order=8, series_block=32,
                                           options=opts)
    feats = ref.preview_features(w, h, z, n, cap, 8, 32, options=opts)
    expect = fav_sched.tile_features(preview_work, w, h, fav_sched.log10_of_zoom(z))
    np.testing.assert_allclose(feats, expect, rtol=1e-6, atol=1e-7)
    assert preview_work.max() <= cap
    ref.close()


def test_mandelbrot_learned_render_is_the_plain_field(gpu_product):
    print("[gpu] mandelbrot learned render 640x360", flush=True)
    lib = fav_hip.load_library()
    ref = c1_reference(lib)
    w, h, z, n = C1["width"], C1["height"], C1["zoom"], C1["max_iter"]
    opts = fav_sched.case_options(C1)
    features, _ = fav_sched.field_features(ref, C1)
    costs = fav_sched.tile_costs(fav_sched.field_work(ref, C1), w, h)
    model = fav_sched.CostModel(4)
    model.train(features, fav_sched.cost_targets(costs), 200, learning_rate=3e-3)
    field, order, timings = fav_sched.render_learned(ref, model, w, h, z, n, options=opts, series=(8, 32))
    plain = ref.render_scheduled(w, h, z, n, series_order=8, series_block=32, options=opts)
    np.testing.assert_array_equal(field.view(np.uint32), plain.view(np.uint32))
    np.testing.assert_array_equal(order, fav_sched.heaviest_first(model.predict(features)))
    assert not np.array_equal(order, np.arange(order.size))
    assert timings["preview_ms"] > 0 and timings["render_ms"] > 0
    ref.close()


def test_batch_in_learned_order_keeps_every_field(gpu_product):
    """12 keyframes (zoom 1 to 1e100, 200x120 and 160x96, caps scaled by
    1/4) in field-major order and in the cost model's order: every field
    bit-identical."""
    print("[gpu] batch learned order, 12 fields", flush=True)
    lib = fav_hip.load_library()
    fields = fav_batch.segment_fields(count=12, seed=5, sizes=((200, 120), (160, 96)), cap_scale=0.25)
    batch = fav_batch.Batch(fields, library=lib)
    try:
        batch.set_order(0)
        batch.render(5)
        plain = [batch.field(i).copy() for i in range(len(fields))]
        x, t = [], []
        for i, f in enumerate(fields):
            case = fav_sched.segment_case(f, f"f{i}")
            feats, _ = fav_sched.field_features(batch.references[i], case)
            x.append(feats)
            t.append(fav_sched.cost_targets(fav_sched.tile_costs(fav_sched.field_work(batch.references[i], case),
                                                                 f["width"], f["height"])))
        model = fav_sched.CostModel(6)
        model.train(np.vstack(x), np.concatenate(t), 200, learning_rate=3e-3)
        costs, times = fav_batch.learned_tile_costs(batch, model)
        assert costs.size == batch.tile_count and np.isfinite(costs).all() and times["preview_ms"] > 0
        assert np.unique(costs).size > 1
        batch.set_order(2, costs)
        batch.render(5)
        for i in range(len(fields)):
            np.testing.assert_array_equal(batch.field(i).view(np.uint32), plain[i].view(np.uint32))
    finally:
        batch.close()
