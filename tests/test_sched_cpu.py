"""Learned tile scheduler on the CPU: the numpy cost-model oracle pinned by
finite differences and closed forms, the feature / cost / order helpers on
hand-built maps, and the C-ABI's argument checks (no device needed)."""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import fav_mlp  # noqa: E402
import fav_sched  # noqa: E402
import fav_workloads  # noqa: E402


def random_problem(seed, n=12):
    rng = np.random.default_rng(seed)
    w = fav_sched.CostModel.initial_weights(seed).astype(np.float64)
    w[fav_mlp.B1:fav_mlp.W2] = rng.normal(0, 0.1, fav_mlp.HIDDEN)
    w[fav_mlp.B2:fav_mlp.W3] = rng.normal(0, 0.1, fav_mlp.HIDDEN)
    x = rng.uniform(0.0, 1.5, (n, fav_mlp.INPUTS))
    t = rng.uniform(0.0, 3.0, n)
    return w, x, t


def test_oracle_gradient_matches_central_differences():
    """Every parameter block's analytic gradient against (L(w+h) - L(w-h)) / 2h
    in float64; a wrong index, sign or missing ReLU gate in any block fails."""
    w, x, t = random_problem(3)
    g, _ = fav_mlp.gradient(w, x, t)
    rng = np.random.default_rng(7)
    blocks = [(fav_mlp.W1, fav_mlp.B1), (fav_mlp.B1, fav_mlp.W2), (fav_mlp.W2, fav_mlp.B2),
              (fav_mlp.B2, fav_mlp.W3), (fav_mlp.W3, fav_mlp.B3), (fav_mlp.B3, fav_mlp.PARAMS)]
    h = 1e-6
    for lo, hi in blocks:
        for p in rng.choice(np.arange(lo, hi), size=min(40, hi - lo), replace=False):
            wp, wm = w.copy(), w.copy()
            wp[p] += h
            wm[p] -= h
            numeric = (fav_mlp.loss(wp, x, t) - fav_mlp.loss(wm, x, t)) / (2 * h)
            assert abs(numeric - g[p]) <= 1e-6 * max(1.0, abs(numeric)), (lo, p, numeric, g[p])


def test_oracle_forward_is_the_layer_formula_on_a_hand_case():
    """One sample, weights chosen so each layer's output is known: x = e_0,
    W1[0][j] = j - 10 (ReLU keeps j > 10), W2 = identity, W3 = ones, b3 = 0.5."""
    w = np.zeros(fav_mlp.PARAMS)
    w1, _, w2, _, w3, _ = fav_mlp.unpack(w)
    w1[0, :] = np.arange(64) - 10.0
    w2[:, :] = np.eye(64)
    w3[:] = 1.0
    w[fav_mlp.B3] = 0.5
    x = np.zeros((1, 65))
    x[0, 0] = 1.0
    y, h1, h2 = fav_mlp.forward(w, x)
    assert y[0] == pytest.approx(sum(j - 10.0 for j in range(11, 64)) + 0.5)
    np.testing.assert_array_equal(h1, h2)


def test_oracle_adam_first_step_is_signed_learning_rate():
    """From zero moments the first Adam step is -lr * g / (|g| + eps')."""
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    w, m, v = fav_mlp.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 0.01, 1)
    np.testing.assert_allclose(w[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert w[3] == 0.0
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_bf16_rounding_known_values():
    # bf16 keeps 7 fraction bits: the ulp of 1.0 is 2^-7; ties go to the even neighbour
    values = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -3.14159265,
                       65504.0, 1e-30, np.inf, -np.inf], dtype=np.float32)
    expect = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, 1.0, -3.140625, 65536.0, 1e-30, np.inf, -np.inf],
                      dtype=np.float32)
    got = fav_mlp.bf16_round(values)
    np.testing.assert_array_equal(got[:7], expect[:7])
    assert got[7] == pytest.approx(1e-30, rel=2 ** -8)
    np.testing.assert_array_equal(got[8:], expect[8:])
    assert np.isnan(fav_mlp.bf16_round(np.array([np.nan], dtype=np.float32)))[0]
    assert (fav_mlp.bf16_round(values).view(np.uint32) & 0xFFFF == 0).all()


def test_tile_costs_take_the_slowest_lane_of_each_wavefront():
    """A 40x36 field: tiles 2x2, ragged; one heavy pixel per 8x8 square sets
    that square's cost."""
    work = np.zeros((36, 40), dtype=np.int32)
    work[0, 0] = 160        # tile 0, square (0,0)
    work[9, 9] = 16         # tile 0, square (1,1)
    work[33, 39] = 32       # tile 3 (ragged corner)
    work[5, 35] = 7
    work[6, 36] = 9         # tile 1: the square (0, 0) of tile 1 holds 7 and 9
    costs = fav_sched.tile_costs(work, 40, 36)
    np.testing.assert_allclose(costs, [(160 + 16) / 16, 9 / 16, 0.0, 32 / 16])


def test_tile_features_patch_layout_and_zoom_column():
    width, height = 64, 40  # tiles 2x2; preview 16x10 (rows 10..15 edge-padded)
    preview = np.arange(160, dtype=np.int32).reshape(10, 16)
    feats = fav_sched.tile_features(preview, width, height, 100.0)
    assert feats.shape == (4, 65) and feats.dtype == np.float32
    scale = fav_sched.FEATURE_SCALE
    np.testing.assert_allclose(feats[1, :8], np.log1p(preview[0, 8:16]) * scale, rtol=1e-6)
    np.testing.assert_allclose(feats[2, :8], np.log1p(preview[8, 0:8]) * scale, rtol=1e-6)
    np.testing.assert_allclose(feats[2, 16:24], np.log1p(preview[9, 0:8]) * scale, rtol=1e-6)  # edge rows
    np.testing.assert_allclose(feats[:, 64], 1.0)


def test_heaviest_first_is_a_stable_permutation():
    order = fav_sched.heaviest_first(np.array([1.0, 5.0, 5.0, 0.5, 3.0]))
    np.testing.assert_array_equal(order, [1, 2, 4, 0, 3])
    assert order.dtype == np.uint32


def test_spearman_matches_scipy():
    from scipy.stats import spearmanr
    rng = np.random.default_rng(1)
    a = rng.integers(0, 20, 300).astype(float)
    b = a + rng.normal(0, 5, 300)
    assert fav_sched.spearman(a, b) == pytest.approx(spearmanr(a, b).statistic, abs=1e-12)


def test_zoom_log10_and_julia_view_centre():
    assert fav_sched.log10_of_zoom("1e100") == 100
    assert fav_sched.log10_of_zoom("2.5e37") == pytest.approx(37.39794, abs=1e-5)
    from decimal import Decimal, localcontext
    field = fav_workloads.julia_field(seed=0)
    with localcontext() as ctx:
        ctx.prec = 380
        x, y = Decimal(field["x"]), Decimal(field["y"])
        cr, ci = (Decimal(v) for v in field["julia"])
        assert abs(x * x - y * y + cr - x) < Decimal("1e-370")  # z^2 + c = z
        assert abs(2 * x * y + ci - y) < Decimal("1e-370")
        assert (1 - x) ** 2 + y * y < x * x + y * y  # the other fixed point is the smaller one


def test_cost_model_abi_rejects_bad_arguments_without_device(product):
    lib = product
    lib.fractal_hip_cost_model_parameter_count.restype = ctypes.c_int
    assert lib.fractal_hip_cost_model_parameter_count() == fav_mlp.PARAMS == 8449
    lib.fractal_last_error.restype = ctypes.c_char_p
    fake = ctypes.c_void_p(16)
    lib.fractal_hip_cost_model_forward.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int, ctypes.c_void_p]
    assert lib.fractal_hip_cost_model_forward(None, fake, 4, fake) == 1
    assert b"null" in lib.fractal_last_error()
    assert lib.fractal_hip_cost_model_forward(fake, fake, 0, fake) == 1
    assert b"sample count" in lib.fractal_last_error()
    lib.fractal_hip_cost_model_train.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                                         ctypes.c_int, ctypes.c_void_p]
    assert lib.fractal_hip_cost_model_train(fake, fake, fake, fake, fake, 8, 0, 1e-3, 1, fake) == 1
    assert b"steps" in lib.fractal_last_error()
    assert lib.fractal_hip_cost_model_train(fake, fake, fake, fake, fake, 8, 10, 1e-3, 0, fake) == 1
    assert b"Adam step" in lib.fractal_last_error()
    assert lib.fractal_hip_cost_model_train(fake, fake, fake, fake, fake, 8, 10, float("nan"), 1, fake) == 1
    assert b"learning rate" in lib.fractal_last_error()


def test_scheduled_render_rejects_bad_orders_on_host(product):
    """A Mandelbrot handle takes a tile order too; orders that are not
    permutations of the 32x32 tiles, and views below zoom 1e6, are refused
    on the host before any launch."""
    import fav_hip
    lib = fav_hip.load_library()
    opts = fav_hip.NativeRenderOptions()
    ref = fav_hip.NativeReference("-0.75", "0.1", "1e7", 2000, 128, options=opts, library=lib)
    for bad in (np.arange(3), np.zeros(4), np.arange(1, 5)):
        with pytest.raises(RuntimeError, match="every 32x32 tile once"):
            ref.render_scheduled(64, 64, "1e7", 2000, tile_order=bad)
    with pytest.raises(RuntimeError, match="zoom >= 1e6"):
        ref.render_scheduled(64, 64, "1e5", 2000, tile_order=np.arange(4))
    with pytest.raises(RuntimeError, match="zoom >= 1e6"):
        ref.preview_features(64, 64, "1e5", 2000, 4096)
    ref.close()


def test_block_order_and_makespan_model():
    """The launch order of a 32x32 tile order over 16x16 blocks, the block
    costs, and the list-schedule makespan on small closed cases."""
    import fav_sched
    # 70x40: 3x2 tiles of 32, 5x3 blocks of 16 (the last column and row partial)
    order = fav_sched.blocks_of_tiles(np.array([5, 0, 1, 2, 3, 4]), 70, 40)
    # tile 5 = (2, 1): blocks (4, 2) only; tile 0 = (0, 0): blocks 0, 1, 5, 6
    assert order.tolist()[:5] == [14, 0, 1, 5, 6]
    assert sorted(order.tolist()) == list(range(15))
    work = np.zeros((40, 70), dtype=np.int32)
    work[3, 2] = 7        # block 0, wave (0, 0)
    work[20, 40] = 11     # block 1 * 5 + 2 = 7
    work[39, 69] = 5      # block 14
    costs = fav_sched.block_costs(work, 70, 40)
    assert costs.shape == (15,) and costs[0] == 7 and costs[7] == 11 and costs[14] == 5 and costs.sum() == 23
    # two machines: row order 3,3,3,4,5 -> {3,3,5} {3,4} = 11; heaviest first
    # 5,4,3,3,3 -> {5,3} {4,3,3} = 10
    jobs = np.array([3.0, 3.0, 3.0, 4.0, 5.0])
    assert fav_sched.makespan(jobs, np.arange(5), 2) == 11.0
    assert fav_sched.makespan(jobs, np.array([4, 3, 0, 1, 2]), 2) == 10.0
    assert fav_sched.makespan(jobs, np.arange(5), 5) == 5.0
    assert fav_sched.makespan(jobs, np.arange(5), 1) == 18.0


def test_batch_costs_spread_over_batch_tiles():
    import fav_batch
    # 70x40: 3x2 tiles of 32, 5x3 batch tiles of 16
    spread = fav_batch.expand_to_blocks(np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0]), 70, 40)
    assert spread.tolist() == [0, 0, 1, 1, 2, 0, 0, 1, 1, 2, 3, 3, 4, 4, 5]
