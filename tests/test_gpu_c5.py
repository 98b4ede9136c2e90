"""The learned tile schedule on the 3840x2160 and 7680x4320 alternate-formula
fields at zoom 1e100, cap 200000 (the tricorn branching junction and the
seed-1 Julia set, c a seeded point 1e-3 off the main cardioid): the field
rendered in the cost model's heaviest-first order is bit-identical to the
plain grid's, and the plain grid matches the oracle on exact-geometry bands
of rows (top, middle, bottom).

Bar: NaN mask, interior set and integer escape index bit-exact; float
field within one ulp (x87 long double in the reference's last logarithm)."""
import numpy as np
import pytest

import fav_sched
import fav_workloads
import golden_io
import native

pytestmark = pytest.mark.gpu

BAND_ROWS = 16


def cases(width, height):
    return [fav_workloads.tricorn_field("branching-junction", width, height, "1e100", 200000),
            fav_workloads.julia_field(seed=1, width=width, height=height, zoom="1e100", max_iter=200000)]


def learned_against_plain(case):
    """(plain field, learned field, learned order) with a model trained for
    200 steps on the field's own preview features and measured tile costs."""
    ref = fav_sched.reference_for(case)
    try:
        w, h, z, n = case["width"], case["height"], case["zoom"], case["max_iter"]
        sample = fav_sched.sample_field(ref, w, h, z, n)
        model = fav_sched.CostModel(1)
        model.train(sample.features, fav_sched.cost_targets(sample.costs), 200, learning_rate=3e-3)
        learned, order, timings = fav_sched.render_learned(ref, model, w, h, z, n)
        plain = ref.render_scheduled(w, h, z, n)
        print(f"[gpu] {case['name']}: plain-grid and learned renders, learned render {timings['render_ms']:.1f} ms",
              flush=True)
        return plain, learned, order
    finally:
        ref.close()


@pytest.mark.parametrize("index", [0, 1])
def test_4k_learned_order_is_the_plain_grid_and_matches_the_oracle(gpu_product, oracle, index):
    case = cases(3840, 2160)[index]
    print(f"[gpu] {case['name']}", flush=True)
    plain, learned, order = learned_against_plain(case)
    tiles = fav_sched.tile_grid(case["width"], case["height"])
    assert order is not None and sorted(order.tolist()) == list(range(tiles[0] * tiles[1]))
    assert not np.array_equal(order, np.arange(order.size))
    np.testing.assert_array_equal(learned.view(np.uint32), plain.view(np.uint32))
    oref = native.OracleReference(oracle, case["x"], case["y"], case["zoom"], case["max_iter"], case["bits"], 8,
                                  formula=case["formula"], julia=case["julia"])
    meta = {"options": {}}
    for y0 in (0, case["height"] // 2 - BAND_ROWS // 2, case["height"] - BAND_ROWS):
        want, _, _ = oref.render_crop(case["width"], case["height"], (0, y0, case["width"], BAND_ROWS), case["zoom"],
                                      case["max_iter"], 8, 32, native.RenderOptions.defaults())
        golden_io.assert_alt_parity({"field": want, "meta": meta}, plain[y0:y0 + BAND_ROWS])


def test_8k_learned_order_is_the_plain_grid(gpu_product, oracle):
    case = cases(7680, 4320)[1]
    print(f"[gpu] {case['name']}", flush=True)
    plain, learned, order = learned_against_plain(case)
    np.testing.assert_array_equal(learned.view(np.uint32), plain.view(np.uint32))
    oref = native.OracleReference(oracle, case["x"], case["y"], case["zoom"], case["max_iter"], case["bits"], 8,
                                  formula=case["formula"], julia=case["julia"])
    y0 = case["height"] // 2 - BAND_ROWS // 2
    want, _, _ = oref.render_crop(case["width"], case["height"], (0, y0, case["width"], BAND_ROWS), case["zoom"],
                                  case["max_iter"], 8, 32, native.RenderOptions.defaults())
    golden_io.assert_alt_parity({"field": want, "meta": {"options": {}}}, plain[y0:y0 + BAND_ROWS])
