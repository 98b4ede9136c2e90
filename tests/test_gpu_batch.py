"""Batched fields: one launch sequence over several references, sizes,
zooms and caps (shallow direct-renderer fields among them) gives every
field bit-for-bit its single-field render, in every tile order and launch
mode."""
import numpy as np
import pytest

import fav_batch
import fav_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(gpu_product):
    # zooms 10^(40k/11): k = 0, 1 are direct-renderer fields in the batch
    fields = fav_batch.segment_fields(count=12, seed=3, decades=40.0, sizes=((96, 54), (160, 90), (200, 120)),
                                      cap_scale=0.25)
    batch = fav_batch.Batch(fields, threads=4)
    yield batch
    batch.close()


def test_batch_fields_equal_single_field_renders(small_batch):
    print(f"[gpu] batch of {len(small_batch.fields)} fields, {small_batch.tile_count} tiles", flush=True)
    small_batch.render_per_field()
    singles = [small_batch.field(i).copy() for i in range(len(small_batch.fields))]
    for order in (0, 1):
        small_batch.set_order(order)
        for graph in (False, True):
            small_batch.render(5, graph=graph)
            for i, expected in enumerate(singles):
                np.testing.assert_array_equal(small_batch.field(i).view(np.uint32), expected.view(np.uint32),
                                              err_msg=f"order {order} graph {graph} field {i}")


def test_batch_cost_order_and_autotune(small_batch):
    print("[gpu] batch autotune", flush=True)
    small_batch.render_per_field()
    singles = [small_batch.field(i).copy() for i in range(len(small_batch.fields))]
    costs = np.random.default_rng(1).random(small_batch.tile_count).astype(np.float32)
    trials, best = fav_batch.autotune(small_batch, costs=costs, repeats=1)
    assert len(trials) == 9 and best in trials
    for i, expected in enumerate(singles):
        np.testing.assert_array_equal(small_batch.field(i).view(np.uint32), expected.view(np.uint32))


def test_batch_graph_replays_repeatedly(small_batch):
    """A captured graph replayed several times, with eager work between
    replays, gives the eager result every time."""
    print("[gpu] batch graph replays", flush=True)
    small_batch.set_order(0)
    small_batch.render(5, graph=False)
    eager = [small_batch.field(i).copy() for i in range(len(small_batch.fields))]
    for _ in range(3):
        small_batch.render(5, graph=True)
        for i, expected in enumerate(eager):
            np.testing.assert_array_equal(small_batch.field(i).view(np.uint32), expected.view(np.uint32))
        small_batch.render_per_field()


def test_floatexp_batch_fields_equal_single_floatexp_renders(small_batch):
    """The batch in floatexp (one persistent kernel over the tile queue):
    every perturbation field is its single-field floatexp render bit for
    bit, every direct field its FP64 render, in the field-major and cost
    orders and as a graph; switching back restores the FP64 fields."""
    print("[gpu] floatexp batch", flush=True)
    small_batch.set_order(0)
    small_batch.set_precision(False)
    small_batch.render(5, graph=False)
    fp64 = [small_batch.field(i).copy() for i in range(len(small_batch.fields))]
    expected = []
    for i, f in enumerate(small_batch.fields):
        if f["zoom_log10"] < 6.0:
            expected.append(fp64[i])
            continue
        field, _ = small_batch.references[i].render_floatexp(f["width"], f["height"], f["zoom"], f["max_iter"],
                                                             options=small_batch.options)
        expected.append(field)
    costs = np.random.default_rng(2).random(small_batch.tile_count).astype(np.float32)
    small_batch.set_precision(True)
    for order in (0, 2):
        small_batch.set_order(order, costs if order == 2 else None)
        for graph in (False, True):
            small_batch.render(5, graph=graph)
            for i, e in enumerate(expected):
                np.testing.assert_array_equal(small_batch.field(i).view(np.uint32), e.view(np.uint32),
                                              err_msg=f"floatexp order {order} graph {graph} field {i}")
    small_batch.set_precision(False)
    small_batch.set_order(0)
    small_batch.render(5, graph=True)
    for i, e in enumerate(fp64):
        np.testing.assert_array_equal(small_batch.field(i).view(np.uint32), e.view(np.uint32))
