"""Learned tile scheduler on the GPU: the cost model's forward, gradient and
Adam kernels against the numpy oracle, deterministic training, and the
scheduled alternate-formula render (any tile order gives the same field
and work map, equal to the plain grid's, which the C oracle checks).

Bars: forward within 1e-5 relative (fp32 sums in another order); gradient
within 1e-4 of the largest component; fields and work maps bit-exact."""
import sys
from pathlib import Path

import numpy as np
import pytest

import golden_io
import fav_hip
import fav_sched
import fav_workloads

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import fav_mlp  # noqa: E402

pytestmark = pytest.mark.gpu


def problem(seed, n):
    rng = np.random.default_rng(seed)
    model = fav_sched.CostModel(seed)
    model.weights[fav_mlp.B1:fav_mlp.W2] = rng.normal(0, 0.1, 64).astype(np.float32)
    model.weights[fav_mlp.B2:fav_mlp.W3] = rng.normal(0, 0.1, 64).astype(np.float32)
    x = rng.uniform(0.0, 1.5, (n, 65)).astype(np.float32)
    t = rng.uniform(0.0, 3.0, n).astype(np.float32)
    return model, x, t


@pytest.mark.parametrize("n", [1, 300, 4099])
def test_cost_model_forward_and_gradient_match_oracle(gpu_product, n):
    print(f"[gpu] cost model forward/gradient n={n}", flush=True)
    model, x, t = problem(11, n)
    w64, x64 = fav_mlp.effective(model.weights, x)
    y_ref, _, _ = fav_mlp.forward(w64, x64)
    y = model.predict(x)
    np.testing.assert_allclose(y, y_ref, rtol=1e-5, atol=1e-5)
    g, loss = model.gradient(x, t)
    g_ref, loss_ref = fav_mlp.gradient(w64, x64, t.astype(np.float64))
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    scale = float(np.abs(g_ref).max())
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-4 * scale)
    for lo, hi in ((fav_mlp.W1, fav_mlp.B1), (fav_mlp.W2, fav_mlp.B2), (fav_mlp.W3, fav_mlp.PARAMS)):
        assert np.abs(g[lo:hi]).max() > 0  # every layer receives gradient


def test_training_matches_oracle_adam_and_is_deterministic(gpu_product):
    print("[gpu] cost model training", flush=True)
    model, x, t = problem(5, 777)
    w0 = model.weights.copy()
    history = model.train(x, t, 1, learning_rate=1e-3)
    w64, x64 = fav_mlp.effective(w0, x)
    g_ref, loss_ref = fav_mlp.gradient(w64, x64, t.astype(np.float64))
    assert history[0] == pytest.approx(loss_ref, rel=1e-5)
    w1, _, _ = fav_mlp.adam_step(w0.astype(np.float64), g_ref, 0.0, 0.0, 1e-3, 1)
    moved = np.abs(g_ref) > 1e-3 * np.abs(g_ref).max()  # sign(g) is stable where |g| is not tiny
    np.testing.assert_allclose(model.weights[moved], w1[moved], rtol=1e-6, atol=1e-7)

    runs = []
    for _ in range(2):
        m = fav_sched.CostModel(5)
        m.weights[:] = w0
        hist = m.train(x, t, 150, learning_rate=3e-3)
        runs.append((m.weights.copy(), hist))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][1][-1] < 0.5 * runs[0][1][0]


JULIA_SMALL = fav_workloads.julia_field(seed=0, width=200, height=120, zoom="1e30", max_iter=3000)


def julia_reference(case, library):
    opts = fav_hip.NativeRenderOptions()
    return fav_hip.NativeReference(case["x"], case["y"], case["zoom"], case["max_iter"], case["bits"],
                                   options=opts, library=library, formula=case["formula"], julia=case["julia"])


def test_tile_order_never_changes_the_field(gpu_product, oracle):
    print("[gpu] scheduled julia 200x120 orders", flush=True)
    case = JULIA_SMALL
    lib = fav_hip.load_library()
    ref = julia_reference(case, lib)
    w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
    plain, plain_work = ref.render_scheduled(w, h, z, n, work=True, series_order=8, series_block=32)
    tiles = fav_sched.tile_grid(w, h)[0] * fav_sched.tile_grid(w, h)[1]
    rng = np.random.default_rng(3)
    for order in (np.arange(tiles), np.arange(tiles)[::-1], rng.permutation(tiles)):
        field, work = ref.render_scheduled(w, h, z, n, tile_order=order, work=True, series_order=8,
                                           series_block=32)
        np.testing.assert_array_equal(field.view(np.uint32), plain.view(np.uint32))
        np.testing.assert_array_equal(work, plain_work)
    assert plain_work.min() >= 1 and plain_work.max() > 10 * plain_work.min()
    oracle_case = dict(name="julia_small", formula=1, x=case["x"], y=case["y"], ref_zoom=z, zoom=z, max_iter=n,
                       bits=case["bits"], width=w, height=h, julia=list(case["julia"]), clone_zoom=None, options={})
    ofield, _ = golden_io.render_alt_oracle(oracle, oracle_case)
    golden_io.assert_alt_parity({"field": ofield, "meta": oracle_case}, plain)
    ref.close()


def test_scheduled_render_rejects_orders_that_are_not_permutations(gpu_product):
    case = JULIA_SMALL
    lib = fav_hip.load_library()
    ref = julia_reference(case, lib)
    w, h = case["width"], case["height"]
    tiles = fav_sched.tile_grid(w, h)[0] * fav_sched.tile_grid(w, h)[1]
    for bad in (np.arange(tiles - 1), np.zeros(tiles), np.arange(1, tiles + 1)):
        with pytest.raises(RuntimeError, match="every 32x32 tile once"):
            ref.render_scheduled(w, h, case["zoom"], case["max_iter"], tile_order=bad)
    with pytest.raises(RuntimeError, match="zoom >= 1e6"):
        ref.render_scheduled(w, h, "1e5", case["max_iter"], tile_order=np.arange(tiles))
    ref.close()


@pytest.mark.parametrize("limit", [16, 1 << 24])
def test_preview_features_are_the_patch_of_the_budgeted_preview(gpu_product, limit):
    """Device features against the numpy restatement on the preview's work
    map; a budget of `limit` passes caps each pixel's work at the limit."""
    print(f"[gpu] preview features limit {limit}", flush=True)
    case = JULIA_SMALL
    lib = fav_hip.load_library()
    ref = julia_reference(case, lib)
    w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
    pw, ph = fav_sched.preview_size(w, h)
    _, preview_work = ref.render_scheduled(pw, ph, z, n, work=True)
    assert preview_work.max() > 16
    feats = ref.preview_features(w, h, z, n, limit)
    expect = fav_sched.tile_features(np.minimum(preview_work, limit), w, h, fav_sched.log10_of_zoom(z))
    np.testing.assert_allclose(feats, expect, rtol=1e-6, atol=1e-7)
    ref.close()


def test_learned_render_is_the_plain_field_in_predicted_order(gpu_product):
    print("[gpu] learned render 200x120", flush=True)
    case = JULIA_SMALL
    lib = fav_hip.load_library()
    ref = julia_reference(case, lib)
    w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
    sample = fav_sched.sample_field(ref, w, h, z, n)
    model = fav_sched.CostModel(2)
    model.train(sample.features, fav_sched.cost_targets(sample.costs), 100, learning_rate=3e-3)
    field, order, timings = fav_sched.render_learned(ref, model, w, h, z, n)
    plain = ref.render_scheduled(w, h, z, n)
    np.testing.assert_array_equal(field.view(np.uint32), plain.view(np.uint32))
    features, _ = fav_sched.preview_features(ref, w, h, z, n)
    np.testing.assert_array_equal(features, sample.features)
    np.testing.assert_array_equal(order, fav_sched.heaviest_first(model.predict(features)))
    assert timings["render_ms"] > 0 and timings["preview_ms"] > 0 and timings["model_ms"] >= 0
    untrained_field, untrained_order, _ = fav_sched.render_learned(ref, fav_sched.CostModel(2), w, h, z, n)
    assert untrained_order is None  # the plain grid is the fallback
    np.testing.assert_array_equal(untrained_field.view(np.uint32), plain.view(np.uint32))
    ref.close()
