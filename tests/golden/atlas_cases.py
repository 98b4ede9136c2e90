"""Golden cases of the raw KFP atlas colour pass
(fractal_atlas_colourise_kfp_raw, renderer.cpp:17969-18265): a parent and a
child scalar field reprojected into one frame (bilinear, interior-
preserving), the child feathered into the centre, then one KFP pass."""
from fav_workloads import atlas_tile as synthetic_tile  # noqa: F401
from kfp_cases import KALLES_DEFAULT


def _case(name, output=(192, 108), parent=(160, 90, 5000), child=(120, 68, 6000), parent_zoom=1.7,
          child_fraction=0.5, child_zoom=1.25, max_iter=6000, centered=False, options=None, stops="kalles"):
    return dict(name=name, output=list(output), parent=list(parent), child=list(child) if child else None,
                parent_zoom=parent_zoom, child_fraction=child_fraction, child_zoom=child_zoom, max_iter=max_iter,
                centered=centered, options=dict(options if options is not None else KALLES_DEFAULT), stops=stops)


CASES = [
    _case("atlas_child_half"),
    _case("atlas_centered", centered=True, options=dict(KALLES_DEFAULT, field_bias=6000.0)),
    _case("atlas_no_child", child=None, child_fraction=0.0, child_zoom=1.0, parent_zoom=2.3),
    _case("atlas_small_child", child_fraction=0.12, child_zoom=1.0, options=dict(color_method=0, slopes=0,
                                                                                 iter_div=1.0), stops="seeded"),
    _case("atlas_large_child_odd", output=(161, 91), child_fraction=0.93, parent_zoom=1.0, child_zoom=3.0,
          options=dict(KALLES_DEFAULT, differences=1)),
]


def case_fields(case):
    """(parent, child or None) float32 fields of a case."""
    pw, ph, pmax = case["parent"]
    parent = synthetic_tile(pw, ph, pmax, seed=21, centered=case["centered"])
    child = None
    if case["child"]:
        cw, ch, cmax = case["child"]
        child = synthetic_tile(cw, ch, cmax, seed=22, centered=case["centered"])
    return parent, child
