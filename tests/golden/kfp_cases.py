"""KFP colour-transfer golden cases: a scalar field (from the field golden
vectors or a seeded synthetic field), a palette, and FractalKfpOptions."""
import numpy as np

# Kalles' default key colours (palettes/kalles-default.kfp "Colors:" line,
# read as RGB triplets in file order) and its transfer settings
# (visualizer.py:430-444 KALLES_DEFAULT_KFP).
KALLES_DEFAULT_STOPS = [(255, 255, 255), (128, 0, 64), (160, 0, 0), (192, 128, 0),
                        (64, 128, 0), (0, 255, 255), (64, 128, 255), (0, 0, 255)]
KALLES_DEFAULT = dict(iter_div=0.01, ratio=360.0, color_method=7, smooth_method=0, smooth=1, power=2.0,
                      slopes=1, slope_power=50.0, slope_ratio=20.0, slope_angle=45.0, differences=3)


def seeded_stops(count=12, seed=5):
    """Seeded uniform RGB key colours."""
    return np.random.default_rng(seed).integers(0, 256, (count, 3)).astype(np.uint8)


def synthetic_field(width=80, height=45, max_iter=5000, seed=11):
    """Smooth ramps with noise, interior blocks, NaN glitches and edge values."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    field = (200.0 + 30.0 * np.sin(x / 5.0) * np.cos(y / 7.0) + 0.8 * x + rng.normal(0, 2.0, (height, width)))
    field = field.astype(np.float32)
    field[10:16, 20:30] = max_iter            # interior block
    field[0, :5] = max_iter                   # interior on the top edge
    field[30, 40] = np.nan                    # glitch sentinels
    field[44, 79] = np.nan
    field[5, 60] = -3.0                       # below zero
    field[40, 3] = max_iter - 0.25            # just below the cap
    return field


CASES = [
    # (name, field source, max_iter, stops kind, options)
    ("kalles_default_c1", "c1_series", 20000, "kalles", dict(KALLES_DEFAULT)),
    ("kalles_default_bla", "bla_1e25", 50000, "kalles", dict(KALLES_DEFAULT)),
    ("plain_c1", "c1_series", 20000, "seeded", dict(color_method=0, slopes=0, smooth=1, iter_div=1.0)),
    ("synthetic_default", "synthetic", 5000, "kalles", dict(KALLES_DEFAULT)),
]
for method in range(12):
    CASES.append((f"method{method}_synthetic", "synthetic", 5000, "seeded",
                  dict(color_method=method, differences=method % 7, slopes=method % 2, iter_div=0.5 + method,
                       color_offset=17.0 * method, interior_color=(10, 200, 30))))
for diff in range(8):
    CASES.append((f"difference{diff}_bla", "bla_1e25", 50000, "kalles",
                  dict(color_method=5 + diff % 4, differences=diff, slopes=1, slope_angle=30.0 * diff,
                       iter_div=0.03)))
CASES += [
    ("flat_inverse", "synthetic", 5000, "seeded", dict(color_method=0, flat=1, inverse_transition=1, iter_div=3.0)),
    ("nosmooth", "synthetic", 5000, "seeded", dict(color_method=1, smooth=0, slopes=1)),
    ("field_bias", "synthetic", 5000, "seeded", dict(color_method=3, field_bias=100.0, slopes=1, differences=2)),
    ("bailout_presets", "synthetic", 5000, "kalles", dict(color_method=9, bailout_radius_preset=2, power=3.0)),
    ("multicolor", "synthetic", 5000, "seeded",
     dict(color_method=0, iter_div=2.0, multi_colors=[(13.0, 0), (-40.0, 1), (7.0, 2), (29.0, 2)])),
    ("multicolor_blend", "bla_1e25", 50000, "kalles",
     dict(color_method=7, differences=3, slopes=1, blend_multi_color=1,
          multi_colors=[(11.0, 0), (5.0, 1), (-80.0, 2)])),
]
