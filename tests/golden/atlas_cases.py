"""Golden cases of the raw KFP atlas colour pass
(fractal_atlas_colourise_kfp_raw, renderer.cpp:17969-18265): a parent and a
child scalar field reprojected into one frame (bilinear, interior-
preserving), the child feathered into the centre, then one KFP pass."""
import numpy as np

from kfp_cases import KALLES_DEFAULT


def synthetic_tile(width, height, max_iter, seed, centered=False):
    """Smooth escape ramps with interior blocks and a few glitch sentinels;
    with `centered`, stored relative to the tile's own cap (value - max_iter)
    as the production atlas encoding does."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    field = (300.0 + 40.0 * np.sin(x / 6.0) * np.cos(y / 5.0) + 1.3 * x - 0.7 * y
             + rng.normal(0, 1.5, (height, width)))
    field[height // 3:height // 3 + 7, width // 2:width // 2 + 9] = max_iter   # interior block
    field[:, :3] = max_iter                                                  # interior left edge
    field[height - 2, width - 5] = np.nan                                    # glitch sentinel
    field = field.astype(np.float64)
    if centered:
        field = field - max_iter
    return field.astype(np.float32)


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
