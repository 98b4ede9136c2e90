"""Parameters beyond the reference's, in the C oracle, pinned by what the
mathematics fixes.

* BLA eps (`fo_reference_create_tuned`, `fractal_hip_create_reference_tuned`):
  2^-38 is the reference's own table (field bit-equal to the reference's
  golden case); at 2^-53 the blocks are shorter (more of them per pixel)
  and the escape indices are at least as close to the BLA-off render's,
  which carry no truncation error.
* Escape radius 1e6 (the oracle's mode 2, R^2 = 1e12; offered by neither the
  reference nor the HIP library): the escape index itself is right (equal
  to a direct arbitrary-precision iteration to |z| > 1e6), but the
  reference forms the loop-top norm |Z+d|^2 as R^2 + (|Z|^2 - R^2) + ...
  (renderer.cpp:8383-8461), whose absolute rounding is ~R^2 * 2^-53 =
  1.1e-4 at this radius -- above the glitch threshold 1e-7 |Z|^2 for every
  |Z| < 33 -- so the strict render of the 640x360-class view at zoom 1e12
  marks every pixel a glitch.  That is why no radius-1e6 mode is offered."""
import mpmath as mp
import numpy as np
import pytest

import golden_io
import native

X = "-0.743643887037158704752191506114774"
Y = "0.131825904205311970493132056385139"


def render(oracle, mode, w, h, zoom, cap, eps=-38, **options):
    ref = native.OracleReference(oracle, X, Y, zoom, cap, 128, 8, escape_radius_mode=mode,
                                 bla_epsilon_exponent=eps)
    opts = native.RenderOptions.defaults(escape_radius_mode=mode, **options)
    field, planes, stats = ref.render(w, h, zoom, cap, 8, 32, opts, planes=True, stats=True)
    return field.reshape(-1), planes, stats


def test_radius_1e6_escape_index_is_right_but_the_norm_loses_the_glitch_test(oracle):
    w, h, zoom, cap = 12, 7, "1e6", 4000
    field, planes, _ = render(oracle, 2, w, h, zoom, cap)
    mp.mp.dps = 60
    cx, cy = mp.mpf(X), mp.mpf(Y)
    scale = mp.mpf("2.8") / mp.mpf(zoom) / h  # project coordinates: 2.8 / zoom over the height
    checked = exact = 0
    for py in range(h):
        for px in range(w):
            c = mp.mpc(cx + (px - (w - 1) / 2) * scale, cy + ((h - 1) / 2 - py) * scale)
            z = mp.mpc(0)
            n = 0
            while n < cap:
                z = z * z + c
                n += 1
                if abs(z) > 1e6:
                    break
            i = py * w + px
            if n < cap and abs(z) > 1e6:
                # FP64 perturbation against the exact orbit: equal, or one
                # step apart where |z_n| lands within rounding of the radius
                assert abs(int(planes.orbit_iteration[i]) + 1 - n) <= 1, (px, py)
                exact += int(planes.orbit_iteration[i]) + 1 == n
                checked += 1
    assert checked > 20 and exact >= 0.95 * checked
    # the same view at zoom 1e12 (the 1e12 workload's centre), strict:
    # radius 2 resolves every pixel, radius 1e6 none
    f2, _, _ = render(oracle, 0, 64, 36, "1e12", 20000)
    f6, _, s6 = render(oracle, 2, 64, 36, "1e12", 20000)
    assert np.isfinite(f2).all()
    assert np.isnan(f6).all() and int(s6[18]) > 0  # RenderStats glitch counter


def test_bla_epsilon_default_and_finer(oracle):
    g = golden_io.load("bla_1e25")
    m = g["meta"]

    def run(eps, **extra):
        ref = native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"],
                                     m["image_series_order"], bla_epsilon_exponent=eps)
        opts = native.RenderOptions.defaults(**dict(m["options"], **extra))
        f, p, s = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"],
                             m["series_block"], opts, planes=True, stats=True)
        return f.reshape(-1), p.orbit_iteration.copy(), s

    f38, i38, s38 = run(-38)
    np.testing.assert_array_equal(f38.view(np.uint32), g["field"].reshape(-1).view(np.uint32))
    f53, i53, s53 = run(-53)
    _, i_off, _ = run(-38, disable_bla=1)
    blocks38, blocks53 = int(s38[3]) + int(s38[4]), int(s53[3]) + int(s53[4])
    assert blocks53 > blocks38
    assert np.sum(i53 != i_off) <= np.sum(i38 != i_off)
    with pytest.raises(RuntimeError):
        native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], 8,
                               bla_epsilon_exponent=-61)
