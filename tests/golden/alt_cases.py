"""Golden cases of the alternate deep formulas (render_alternate_reference_impl):
Julia (1), burning ship (2) and tricorn (3) at zooms >= 1e6, rendered
through fractal_create_reference_ex_options + fractal_render_reference_ex
on the reference's scalar backend (no planes)."""

# Boundary points found by zooming into escape-count gradients of the
# shallow renderer (burning ship from -1.76-0.03i, tricorn from -0.3+0.6i).
BS_X = "-0.994372265592592641165023394346757480956579524057882803457758"
BS_Y = "0.0494499089999999949595932413615173572907224297523498535156249"
BS_DEEP_X = "-0.99437223817162869336292894990231303651213507961343835901331355555555555555555556"
BS_DEEP_Y = "0.049449897048211376543676574694850690624055763085683186848958233333333333333333334"
TC_X = "-0.000513548037037056036851581755745091531598181636245162398727164"
TC_Y = "0.294643466777777797149998000807929404092849128776126437717014"
# Repelling fixed point (1 + sqrt(1 - 4c))/2 of z^2 + c for c = -0.8 + 0.156i, 70 digits.
JULIA_C = ("-0.8", "0.156")
JULIA_X = "1.527503118643534610789746402444915337566745947811708160339875197003203"
JULIA_Y = "-0.07591217835228786707181419482634804636642219484797802253973259344918689"


def _case(name, formula, x, y, zoom, max_iter, bits=200, width=96, height=54, julia=("0", "0"), clone_zoom=None,
          **options):
    return dict(name=name, formula=formula, x=x, y=y, ref_zoom=zoom, zoom=clone_zoom or zoom, max_iter=max_iter,
                bits=bits, width=width, height=height, julia=list(julia), clone_zoom=clone_zoom, options=options)


def _c3_centre(zoom_exponent):
    """A burning-ship near-boundary centre of c3_centres.json."""
    import json
    from pathlib import Path
    entry = json.loads((Path(__file__).resolve().parent / "c3_centres.json").read_text())[f"2_{zoom_exponent}"]
    return entry["x"], entry["y"]


CASES = [
    _case("alt_burning_ship_1e7", 2, BS_X, BS_Y, "1e7", 4000),
    _case("alt_burning_ship_nobla", 2, BS_X, BS_Y, "1e7", 4000, width=64, height=36, disable_bla=1),
    _case("alt_burning_ship_block16", 2, BS_X, BS_Y, "3e7", 4000, max_linear_bla_length=16),
    _case("alt_burning_ship_1e15", 2, BS_DEEP_X, BS_DEEP_Y, "1e15", 9000, bits=256, width=64, height=36),
    _case("alt_tricorn_1e7", 3, TC_X, TC_Y, "1e7", 4000),
    _case("alt_tricorn_kalles_cycle", 3, TC_X, TC_Y, "2e7", 5000, escape_radius_mode=1, coordinate_mode=1,
          strict=0),
    _case("alt_julia_1e8", 1, JULIA_X, JULIA_Y, "1e8", 4000, julia=JULIA_C),
    _case("alt_julia_kalles_1e8", 1, JULIA_X, JULIA_Y, "1e8", 4000, julia=JULIA_C, escape_radius_mode=1,
          coordinate_mode=1),
    _case("alt_julia_clone", 1, JULIA_X, JULIA_Y, "1e8", 4000, julia=JULIA_C, clone_zoom="1e11"),
    # near-boundary precision centres (burning-ship period-68 / period-226
    # nucleus plus a 1e-32 / 1e-102 offset), the 2560x1440 fields' views at
    # 64x36
    _case("c3_burning_ship_1e30", 2, *_c3_centre(30), "1e30", 100000, bits=192, width=64, height=36,
          allow_recovery=1),
    _case("c3_burning_ship_1e100", 2, *_c3_centre(100), "1e100", 100000, bits=416, width=64, height=36,
          allow_recovery=1),
]
