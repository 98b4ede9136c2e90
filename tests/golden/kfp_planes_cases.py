"""Golden cases of the plane-aware KFP colour pass
(fractal_colourise_kfp_planes with Kalles planes, renderer.cpp:16314-16545,
6696-7289, 5655-5976).  The planes come from the field golden vectors'
render planes (orbit iteration, phase, DE, |z|^2 tests); stored iteration /
transition, bailout and texture planes are derived from them with a seed
(tests/kfp_planes_io.py)."""
from kfp_cases import KALLES_DEFAULT

FULL = ("orbit_iteration", "phase", "de_x", "de_y", "test1", "test2")


def _c(name, source, max_iter, planes, options, stops="kalles", seed=1, **extra):
    return dict(name=name, source=source, max_iter=max_iter, planes=planes, options=options, stops=stops, seed=seed,
                **extra)


CASES = [
    # Kalles default from the full orbit planes: smoothing from test1, stencil differences, slopes
    _c("kfp_planes_default_c1", "c1_series", 20000, FULL, dict(KALLES_DEFAULT)),
    _c("kfp_planes_default_bla", "bla_1e25", 50000, FULL, dict(KALLES_DEFAULT)),
    # analytic DE differences with relief
    _c("kfp_planes_analytic_de", "c1_series", 20000, FULL,
       dict(color_method=5, differences=7, slopes=1, slope_power=40.0, slope_ratio=30.0, slope_angle=60.0,
            iter_div=0.05, smooth=1)),
    # phase colour strength on the plain transfer
    _c("kfp_planes_phase", "bla_1e25", 50000, FULL,
       dict(color_method=0, iter_div=2.0, phase_color_strength=40.0, smooth=1), stops="seeded"),
    # the test2 smoothing (norm preset 1) and the method-4 range over the plane samples
    _c("kfp_planes_test2_stretch", "c1_kalles", 20000, FULL,
       dict(color_method=4, smooth_method=1, bailout_norm_preset=1, slopes=1, differences=2), stops="seeded"),
    # stored iteration + transition planes (no orbit plane), hidden glitches
    _c("kfp_planes_iteration_hidden_glitches", "c1_series", 20000, ("iteration",),
       dict(color_method=1, differences=1, slopes=1, show_glitches=0), transition_glitches=40),
    # per-pixel bailout plane with the test1 smoothing
    _c("kfp_planes_bailout", "bla_1e27", 50000, FULL + ("bailout",),
       dict(color_method=3, slopes=0, iter_div=0.3), stops="seeded", seed=2),
    # texture merged into the colour (interior textured too)
    _c("kfp_planes_texture", "c1_series", 20000, FULL,
       dict(color_method=7, differences=3, texture_enabled=1, texture_merge=0.4, texture_power=8.0,
            texture_ratio=30.0, iter_div=0.01, slopes=1, slope_power=50.0, slope_ratio=20.0), texture=True),
    # a planes struct carrying only phase: the scalar field path plus phase, method-4 range via the plane path
    _c("kfp_planes_phase_only", "bla_1e16_series", 30000, ("phase",),
       dict(color_method=4, phase_color_strength=-25.0, differences=4, slopes=1), stops="seeded"),
]


def _a(name, parent, child, ow, oh, parent_zoom, fraction, child_zoom, max_iter, options, stops="kalles", seed=1,
       **extra):
    return dict(name=name, parent=parent, child=child, ow=ow, oh=oh, parent_zoom=parent_zoom, fraction=fraction,
                child_zoom=child_zoom, max_iter=max_iter, options=options, stops=stops, seed=seed, **extra)


# fractal_atlas_colourise_kfp_planes (renderer.cpp:17323-17801): tiles are
# field golden vectors with their render planes
ATLAS_CASES = [
    _a("kfp_atlas_planes_child_third", "c1_series", "c1_noseries", 120, 68, 1.5, 0.33, 1.2, 20000,
       dict(KALLES_DEFAULT)),
    _a("kfp_atlas_planes_centred", "c1_series", "c1_noseries", 120, 68, 1.3, 0.5, 1.1, 20000,
       dict(KALLES_DEFAULT, field_bias=20000.0), centred=1),
    _a("kfp_atlas_planes_full_child", "c1_series", "c1_kalles", 100, 56, 1.2, 1.0, 1.4, 20000,
       dict(color_method=4, smooth_method=1, bailout_norm_preset=1, slopes=1, differences=2), stops="seeded"),
    _a("kfp_atlas_planes_no_child", "bla_1e25", None, 80, 45, 1.6, 0.0, 1.0, 50000,
       dict(color_method=5, differences=7, slopes=1, iter_div=0.05)),
    # the child lacks phase, carries hidden glitches; the parent carries phase
    _a("kfp_atlas_planes_mixed", "c1_series", "c1_noseries", 128, 72, 1.4, 0.45, 1.3, 20000,
       dict(color_method=0, iter_div=2.0, phase_color_strength=40.0, show_glitches=0, slopes=1, differences=3),
       stops="seeded", child_planes=("orbit_iteration", "de_x", "de_y", "test1", "test2"),
       child_transition_glitches=30),
    # a texture on the child tile only
    _a("kfp_atlas_planes_child_texture", "c1_series", "c1_noseries", 96, 54, 1.2, 0.6, 1.0, 20000,
       dict(KALLES_DEFAULT, texture_enabled=1, texture_merge=0.5, texture_power=4.0, texture_ratio=50.0),
       child_texture=True, parent_planes=("orbit_iteration", "test1", "test2")),
]


# fractal_crop_colourise_kfp (renderer.cpp:17803-17968): one scalar tile,
# optionally centred on its cap (then field_bias is the cap)
CROP_CASES = [
    dict(name="kfp_crop_default", source="c1_series", ow=80, oh=45, zoom=1.6, max_iter=20000,
         options=dict(KALLES_DEFAULT), stops="kalles", centred=False),
    dict(name="kfp_crop_stretch", source="bla_1e25", ow=96, oh=54, zoom=2.0, max_iter=50000,
         options=dict(color_method=4, differences=1, slopes=1), stops="seeded", centred=False),
    dict(name="kfp_crop_centred", source="c1_kalles", ow=70, oh=40, zoom=1.3, max_iter=20000,
         options=dict(KALLES_DEFAULT, field_bias=20000.0), stops="kalles", centred=True),
]


# fractal_atlas_colourise_kfp (renderer.cpp:16750-17321): the output-sized
# parent and a child field at an explicit rectangle
RECT_CASES = [
    dict(name="kfp_rect_feather", parent="c1_series", child="bla_1e27", left=20, top=12, feather=8, max_iter=50000,
         options=dict(KALLES_DEFAULT), stops="kalles"),
    dict(name="kfp_rect_stretch", parent="c1_series", child="c2_kalles", left=40, top=20, feather=12, max_iter=100000,
         options=dict(color_method=4, differences=2, slopes=1), stops="seeded"),
    dict(name="kfp_rect_hard_edge", parent="bla_1e25", child="bla_1e27", left=8, top=5, feather=0, max_iter=50000,
         options=dict(color_method=7, differences=3, slopes=1, iter_div=0.02), stops="kalles"),
    dict(name="kfp_rect_no_child", parent="c1_kalles", child=None, left=0, top=0, feather=0, max_iter=20000,
         options=dict(KALLES_DEFAULT), stops="kalles"),
]
