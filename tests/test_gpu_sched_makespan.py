"""What the learned schedule gains, without timing noise: a cost model is
trained on the GPU (deterministic kernels) over the seeded training fields
(the Mandelbrot-map families at 640x360 zoom 1e12, 1920x1080 zoom 1e100,
2560x1440 near-nucleus fields and atlas keyframes, and the Julia / tricorn
fields at 960x540), and on held-out fields the makespan of the learned
heaviest-first order is computed from each field's measured per-pixel work
map (fav_sched.makespan: blocks of 16x16 pixels, each as long as its
slowest lane, handed to the first free of the kernel's resident block
slots) against the plain grid's.

The numbers are a model of the dispatcher, not a timing: every input
(work maps, preview features, trained weights, order) is bit-reproducible,
so the result is too."""
import numpy as np
import pytest

import fav_sched
import fav_workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(gpu_product):
    print("[gpu] makespan: training the cost model", flush=True)
    fields = fav_sched.mandelbrot_training_fields() + fav_sched.training_fields(width=960, height=540)
    m, summary = fav_sched.train_model(fields, schedule=((3e-3, 1000), (1e-3, 1000)))
    print(f"[gpu] makespan: trained on {summary['samples']} tiles, loss {summary['loss_first']:.2f} -> "
          f"{summary['loss_last']:.3f}", flush=True)
    return m


# the learned-schedule configuration (3840x2160, zoom 1e100, cap 200000:
# Julia seeds 1, 2, 4 and the tricorn branching junction) and three
# Mandelbrot-map fields: the Misiurewicz spiral, a near-nucleus view and the
# all-interior zoom-1e100 field
HELD_OUT = [(f"julia seed {s}", fav_workloads.julia_field(seed=s, width=3840, height=2160, zoom="1e100",
                                                          max_iter=200000)) for s in (1, 2, 4)] + [
    ("tricorn junction", fav_workloads.tricorn_field("branching-junction", 3840, 2160, "1e100", 200000)),
    ("misiurewicz", dict(fav_workloads.misiurewicz_field(), formula=0)),
    ("nucleus 1e30", fav_workloads.near_boundary_field(0, 1, 30)),
    ("interior 1e100", dict(fav_workloads.field_1e100_1080p(), formula=0)),
]


def makespans(model, case):
    ref = fav_sched.reference_for(case)
    try:
        work = fav_sched.field_work(ref, case)
        features, _ = fav_sched.field_features(ref, case)
    finally:
        ref.close()
    return fav_sched.makespan_report(work, model.predict(features), case["width"], case["height"],
                                     fav_sched.kernel_slots(case.get("formula", 0)))


def test_learned_order_shortens_the_makespan(model):
    rows = {}
    for name, case in HELD_OUT:
        print(f"[gpu] makespan: {case['name']}", flush=True)
        r = makespans(model, case)
        rows[name] = r
        print(f"  {name:18s} plain {r['plain']:10.0f}  learned {r['learned']:10.0f}  measured-cost order "
              f"{r['measured_order']:10.0f}  bound {r['bound']:10.0f}  gain {r['gain_pct']:5.1f}%", flush=True)
    # an all-interior field has no tail to remove: every order is equal
    assert rows["interior 1e100"]["learned"] == rows["interior 1e100"]["plain"]
    for name, r in rows.items():
        assert r["learned"] >= r["bound"] - 1e-9  # the bound holds for any order
        assert r["learned"] <= 1.05 * r["plain"], name  # never much worse than the plain grid
    long_tailed = [rows[n]["gain_pct"] for n in ("julia seed 1", "julia seed 2", "julia seed 4", "tricorn junction",
                                                 "misiurewicz")]
    assert np.mean(long_tailed) >= 3.0
    assert max(long_tailed) >= 10.0
