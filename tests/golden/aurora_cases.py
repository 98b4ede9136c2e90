"""Golden cases for the ordinary-palette colouriser and the per-frame crop /
atlas compositors (renderer.h:400-449, 713-835, 943-1068).  Each case names
one function, its geometry and controls and an input seed
(tests/aurora_io.py regenerates the inputs); ``make_golden_aurora.py``
records the output of the reference's renderer.cpp for it."""

ACCENTS = [250, 40, 90, 30, 200, 120, 60, 80, 240]
COLOUR = dict(phase=1.3, vocal=0.7, instrumental=0.2, pitch=0.8)
NEUTRAL = dict(phase=-2.1, vocal=0.0, instrumental=0.9, pitch=0.5)  # no hue rotation


def _c(name, kind, fn, seed, **kw):
    return dict(name=name, kind=kind, fn=fn, seed=seed, **kw)


CASES = [
    _c("aurora_colourise", "colourise", "fractal_colourise", 1, sw=96, sh=54, max_iter=5000, **COLOUR),
    _c("aurora_colourise_small_cap", "colourise", "fractal_colourise", 2, sw=64, sh=36, max_iter=800, **NEUTRAL),
    _c("aurora_colourise_accents", "colourise", "fractal_colourise_accents", 3, sw=96, sh=54, max_iter=20000,
       accents=ACCENTS, interior=(12, 34, 56), **COLOUR),
    _c("aurora_apply_accents", "accents", "fractal_apply_aurora_accents", 4, w=80, h=45, accents=ACCENTS, pitch=0.15),
    _c("aurora_apply_accents_neutral", "accents", "fractal_apply_aurora_accents", 5, w=40, h=30, accents=ACCENTS,
       pitch=0.5),
    _c("aurora_crop_rgb", "crop_rgb", "fractal_crop_rgb", 6, sw=96, sh=54, ow=80, oh=45, zoom=1.7),
    _c("aurora_crop_rgb_identity", "crop_rgb", "fractal_crop_rgb", 7, sw=48, sh=27, ow=48, oh=27, zoom=0.5),
    _c("aurora_atlas_rgb", "atlas_rgb", "fractal_atlas_composite_rgb", 8, pw=96, ph=54, cw=96, ch=54, ow=120, oh=68,
       parent_zoom=1.3, fraction=0.4, child_zoom=1.2, feather=12),
    _c("aurora_atlas_rgb_no_child", "atlas_rgb", "fractal_atlas_composite_rgb", 9, pw=96, ph=54, cw=0, ch=0, ow=64,
       oh=36, parent_zoom=2.0, fraction=0.0, child_zoom=1.0, feather=0, child=False),
    _c("aurora_crop_field", "crop_field", "fractal_crop_field", 10, sw=96, sh=54, ow=70, oh=40, zoom=2.2,
       max_iter=3000),
    _c("aurora_atlas_field", "atlas_field", "fractal_atlas_field", 11, pw=96, ph=54, pmax=3000, cw=96, ch=54,
       cmax=4000, ow=128, oh=72, parent_zoom=1.5, fraction=0.5, palette_max_iter=3500),
    _c("aurora_atlas_field_ex_centred", "atlas_field", "fractal_atlas_field_ex", 12, pw=96, ph=54, pmax=3000,
       cw=80, ch=45, cmax=4000, ow=128, oh=72, parent_zoom=1.25, fraction=0.35, child_zoom=1.4, pbias=3000.0,
       cbias=4000.0, obias=4000.0, palette_max_iter=2000, feather=20),
    _c("aurora_atlas_field_ex_full_child", "atlas_field", "fractal_atlas_field_ex", 13, pw=64, ph=36, pmax=3000,
       cw=96, ch=54, cmax=4000, ow=100, oh=56, parent_zoom=1.1, fraction=1.0, child_zoom=1.3, palette_max_iter=1000,
       feather=8),
    _c("aurora_atlas_field_ex_no_child", "atlas_field", "fractal_atlas_field_ex", 14, pw=96, ph=54, pmax=3000,
       cw=0, ch=0, cmax=0, ow=60, oh=34, parent_zoom=1.7, fraction=0.0, child_zoom=1.0, palette_max_iter=3000,
       feather=8, child=False),
    _c("aurora_crop_colourise", "crop_colourise", "fractal_crop_colourise", 15, sw=96, sh=54, ow=80, oh=45,
       zoom=1.6, max_iter=6000, **COLOUR),
    _c("aurora_crop_colourise_interior", "crop_colourise", "fractal_crop_colourise_interior", 16, sw=96, sh=54,
       ow=96, oh=54, zoom=1.0, max_iter=900, interior=(200, 10, 30), **NEUTRAL),
    _c("aurora_crop_colourise_accents", "crop_colourise", "fractal_crop_colourise_accents", 17, sw=96, sh=54, ow=72,
       oh=40, zoom=2.5, max_iter=6000, accents=ACCENTS, interior=(1, 2, 3), **COLOUR),
    _c("aurora_atlas_colourise", "atlas_colourise", "fractal_atlas_colourise", 18, pw=96, ph=54, pmax=3000, cw=96,
       ch=54, cmax=5000, ow=160, oh=90, parent_zoom=1.4, fraction=0.6, palette_max_iter=0, **COLOUR),
    _c("aurora_atlas_colourise_interior_full", "atlas_colourise", "fractal_atlas_colourise_interior", 19, pw=64,
       ph=36, pmax=3000, cw=96, ch=54, cmax=5000, ow=120, oh=68, parent_zoom=1.2, fraction=1.0,
       palette_max_iter=7000, interior=(9, 99, 199), **NEUTRAL),
    _c("aurora_atlas_colourise_accents", "atlas_colourise", "fractal_atlas_colourise_accents", 20, pw=96, ph=54,
       pmax=4000, cw=80, ch=45, cmax=4000, ow=200, oh=112, parent_zoom=1.8, fraction=0.45, palette_max_iter=4500,
       accents=ACCENTS, interior=(40, 50, 60), **COLOUR),
    _c("aurora_atlas_colourise_no_child", "atlas_colourise", "fractal_atlas_colourise", 21, pw=96, ph=54, pmax=3000,
       cw=0, ch=0, cmax=0, ow=64, oh=36, parent_zoom=1.0, fraction=0.0, palette_max_iter=3000, child=False,
       **NEUTRAL),
]
