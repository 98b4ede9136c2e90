"""Golden cases for the glitch-repair driver (fav_repair.repair_glitches,
the region queue of visualizer.py:8215-9252).  Each case is a strict field
render whose NaN pixels the driver repairs with secondary references and
point renders."""
from cases import X100, Y100

# The catalogue's "Old Wood Dish" (deep_zoom_points.py:85-92), the point of the
# reference's own glitch-repair test (tests/test_visualizer.py:3393-3440).
OLD_WOOD_X = ("-7.6211793767976386831524124703027542346975542662773842018031824527079880488711666681577392689191400914"
              "23417335237629214081127371665327888360528335010378835950994472786588199757e-1")
OLD_WOOD_Y = ("-9.5731755938079152181101271958769817291852842957531737800785254979881669321686600681706314310158846138"
              "50090347745780605812248444859260422571076360932193737688403082995582081506e-2")


def _case(name, x, y, log10_zoom, width, height, max_iter, series_order=3, series_block=256, **initial):
    return dict(name=name, x=x, y=y, log10_zoom=log10_zoom, width=width, height=height, max_iter=max_iter,
                series_order=series_order, series_block=series_block, initial=initial)


CASES = [
    # natural cancellation glitches of a strict render: two small regions
    _case("repair_old_wood_dish_1e60", OLD_WOOD_X, OLD_WOOD_Y, 60.0, 320, 180, 4000, strict_cycle=1),
    # BLA off at 1e100: every pixel is a strict glitch, one 9216-pixel region
    # that the driver splits into quadrants before the point renders
    _case("repair_split_1e100", X100, Y100, 100.0, 96, 96, 20000, disable_bla=1, series_max_terms=8),
]
