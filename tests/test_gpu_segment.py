"""One atlas segment at full size: 48 keyframe fields from zoom 1 to 1e100
(the first three below 1e6, rendered by the direct double renderer),
sizes drawn from 1280x720, 1920x1080 and 3840x2160, caps 2000 + 1000 per
decade up to 100000, each with its own reference.  The batched launch
gives every field bit-for-bit its single-field render, and one field per
size class (plus one direct field) matches the oracle on an
exact-geometry band of rows through its centre.

Bar: NaN mask, integer escape index bit-exact; float field within one ulp
(x87 long double in the reference's last logarithm)."""
import numpy as np
import pytest

import fav_batch
import golden_io
import native

pytestmark = pytest.mark.gpu

BAND_ROWS = 8


@pytest.fixture(scope="module")
def segment(gpu_product):
    fields = fav_batch.segment_fields()
    print(f"[gpu] segment of {len(fields)} fields, {sum(f['width'] * f['height'] for f in fields) / 1e6:.1f} Mpx",
          flush=True)
    batch = fav_batch.Batch(fields, threads=16)
    yield batch
    batch.close()


def assert_same_field(got, want, label):
    got, want = got.reshape(-1), want.reshape(-1)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want), err_msg=label)
    ok = ~np.isnan(want)
    ulps = np.abs(got[ok].view(np.int32).astype(np.int64) - want[ok].view(np.int32).astype(np.int64))
    assert ulps.size == 0 or int(ulps.max()) <= golden_io.FIELD_ULPS, f"{label}: {int(ulps.max())} ulp"
    np.testing.assert_array_equal(np.floor(got[ok].astype(np.float64)), np.floor(want[ok].astype(np.float64)),
                                  err_msg=label)


def test_segment_batch_equals_per_field_renders(segment):
    segment.render_per_field()
    singles = [segment.field(i).copy() for i in range(len(segment.fields))]
    for graph in (True, False):
        segment.set_order(1)
        ms = segment.render(5, graph=graph)
        print(f"[gpu] segment batch (graph {graph}): {ms:.1f} ms", flush=True)
        for i, expected in enumerate(singles):
            np.testing.assert_array_equal(segment.field(i).view(np.uint32), expected.view(np.uint32),
                                          err_msg=f"graph {graph} field {i}")


def chosen_fields(fields):
    """The deepest field of each size class and the deepest direct field."""
    picks = {}
    for i, f in enumerate(fields):
        key = "direct" if f["zoom_log10"] < 6.0 else (f["width"], f["height"])
        picks[key] = i
    return sorted(picks.values())


def test_segment_fields_match_the_oracle(segment, oracle):
    segment.render(5, graph=True)
    opts = native.RenderOptions.defaults(**{k: getattr(segment.options, k) for k in
                                            ("series_min_terms", "escape_radius_mode", "coordinate_mode")})
    for i in chosen_fields(segment.fields):
        f = segment.fields[i]
        got = segment.field(i)
        y0 = f["height"] // 2 - BAND_ROWS // 2
        print(f"[gpu] segment field {i}: {f['width']}x{f['height']} zoom {f['zoom']} cap {f['max_iter']}",
              flush=True)
        if f["zoom_log10"] < 6.0:
            case = dict(width=f["width"], height=f["height"], zoom=f["zoom"], x=f["x"], y=f["y"],
                        max_iter=f["max_iter"], formula=0, planes=False,
                        options=dict(escape_radius_mode=1, coordinate_mode=1))
            want, _ = native.render_direct_oracle(oracle, case)
            assert_same_field(got, want, f"direct field {i}")
            continue
        ref = native.OracleReference(oracle, f["x"], f["y"], f["zoom"], f["max_iter"], f["bits"], 8,
                                     escape_radius_mode=1, coordinate_mode=1)
        want, _, _ = ref.render_crop(f["width"], f["height"], (0, y0, f["width"], BAND_ROWS), f["zoom"],
                                     f["max_iter"], 3, 256, opts)
        assert_same_field(got[y0:y0 + BAND_ROWS], want, f"field {i} rows {y0}+{BAND_ROWS}")
