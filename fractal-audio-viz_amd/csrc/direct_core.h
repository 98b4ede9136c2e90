// direct_core.h -- the shallow direct renderer: z -> f(z) + c in plain
// doubles per pixel (render_direct_scalar_formula /
// render_direct_with_planes_formula, renderer.cpp:10054-10136, 10280-10378).
//
// Formulas: 0 Mandelbrot, 1 Julia (z0 = pixel, c = constant), 2 burning
// ship (|x|, |y| folded before squaring), 3 tricorn (conjugate).  Every
// product and sum is rounded in the reference's order (-ffp-contract=off),
// so escape indices are bit-identical.  Compiled for the gfx950 kernel
// (direct_kernel.hip) and, unchanged, for the test-only CPU harness.
#pragma once

#include "field_core.h"

namespace fav {
namespace direct {

constexpr uint32_t kHintSkipPhase = 1U;       // FRACTAL_RENDER_HINT_SKIP_PHASE
constexpr uint32_t kHintSkipAnalyticDe = 2U;  // FRACTAL_RENDER_HINT_SKIP_ANALYTIC_DE

struct DirectParams {
    const double* x;  // per-column real coordinate (width entries)
    const double* y;  // per-row imaginary coordinate (height entries)
    float* output;
    int64_t* plane_iteration;  // all plane pointers null when not requested
    double *plane_phase, *plane_de_x, *plane_de_y, *plane_test1, *plane_test2;
    int32_t width, height, max_iter, formula;
    uint32_t plane_flags;
    int32_t pad;
    double julia_re, julia_im, output_bias;
    double escape2;       // R^2: 4 or 1e8
    double safe_escape2;  // R^2 + 1e-7
    double log_power;     // log(2)
    double spacing_m;     // pixel spacing as FloatExp (DE scale)
    int32_t spacing_e;
};

// One orbit step of formula F (iterate_direct_formula_static /
// iterate_direct_formula_with_derivative, renderer.cpp:251-398).  The
// derivative is carried for Mandelbrot (d/dc, +1) and Julia (d/dz0, +0).
template <int F, bool Derivative>
FAV_HD void step(double& zr, double& zi, double pr, double pi, double& dr, double& di) {
    double nr, ni;
    if constexpr (F == 2) {
        const double ar = fabs_(zr), ai = fabs_(zi);
        const double rs = ar * ar, is = ai * ai, cross = 2.0 * ar * ai;
        nr = (rs - is) + pr;
        ni = cross + pi;
    } else if constexpr (F == 3) {
        const double rs = zr * zr, is = zi * zi, cross = -2.0 * zr * zi;
        nr = (rs - is) + pr;
        ni = cross + pi;
    } else {
        nr = zr * zr - zi * zi + pr;
        ni = 2.0 * zr * zi + pi;
    }
    if constexpr (Derivative) {
        const double add_real = F == 1 ? 0.0 : 1.0;
        const double ndr = 2.0 * zr * dr - 2.0 * zi * di + add_real;
        const double ndi = 2.0 * zr * di + 2.0 * zi * dr + 0.0;
        dr = ndr;
        di = ndi;
    }
    zr = nr;
    zi = ni;
}

// (iteration + 1) - log(log|z|)/log 2 - bias, rounded to float.  The
// reference forms the subtraction in x87 long double; the device forms it
// in double, which can differ by one float ulp at a rounding boundary.
FAV_HD float smooth_value(int iteration, double magnitude, double log_power, double bias) {
    const double q = ::log(::log(magnitude)) / log_power;
#if defined(__HIP_DEVICE_COMPILE__)
    return static_cast<float>(static_cast<double>(iteration + 1) - q - bias);
#else
    return static_cast<float>((static_cast<long double>(iteration + 1) - q) - static_cast<long double>(bias));
#endif
}

template <int F, bool Planes, bool Derivative>
FAV_HD void pixel(const DirectParams& p, int px, int py) {
    const double cx = p.x[px], cy = p.y[py];
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    const core::PlaneOut planes{p.plane_iteration, p.plane_phase, p.plane_de_x, p.plane_de_y, p.plane_test1,
                                p.plane_test2};
    if constexpr (F == 0) {  // main cardioid and period-2 bulb, renderer.cpp:10087-10097
        const double q = (cx - 0.25) * (cx - 0.25) + cy * cy;
        const bool in_cardioid = q * (q + cx - 0.25) <= 0.25 * cy * cy;
        const bool in_bulb = (cx + 1.0) * (cx + 1.0) + cy * cy <= 0.0625;
        if (in_cardioid || in_bulb) {
            p.output[index] = core::encode_iteration(p.max_iter, p.output_bias);
            if constexpr (Planes) core::clear_planes(planes, index, p.max_iter);
            return;
        }
    }
    constexpr bool julia = F == 1;
    double zr = julia ? cx : 0.0, zi = julia ? cy : 0.0;
    const double pr = julia ? p.julia_re : cx, pi = julia ? p.julia_im : cy;
    double dr = julia ? 1.0 : 0.0, di = 0.0;
    double test1 = 0.0, test2 = 0.0;
    for (int iteration = 0; iteration < p.max_iter; ++iteration) {
        step<F, Derivative>(zr, zi, pr, pi, dr, di);
        const double m2 = zr * zr + zi * zi;
        test2 = test1;
        test1 = m2;
        if (m2 > p.escape2 || !isfinite_(m2)) {
            const double safe = isfinite_(m2) ? (m2 < p.safe_escape2 ? p.safe_escape2 : m2) : 1.7976931348623157e308;
            p.output[index] = smooth_value(iteration, ::sqrt(safe), p.log_power, p.output_bias);
            if constexpr (Planes) {
                planes.iteration[index] = iteration;  // Kalles antal of the first escaped sample
                double phase = 0.0;
                if ((p.plane_flags & kHintSkipPhase) == 0U && isfinite_(zr) && isfinite_(zi)) {
                    phase = ::atan2(zi, zr) / core::kTwoPi;
                    phase -= ::floor(phase);
                }
                planes.phase[index] = phase;
                double dx = 0.0, dy = 0.0;
                if constexpr (Derivative) {
                    if (isfinite_(zr) && isfinite_(zi) && isfinite_(dr) && isfinite_(di)) {
                        const SC total = sc_from_fe(fe_from_parts(zr, 0), fe_from_parts(zi, 0));
                        const SC dz = sc_from_fe(fe_from_parts(dr, 0), fe_from_parts(di, 0));
                        // the reference keeps whatever render_plane_de wrote,
                        // non-finite results included (renderer.cpp:10412-10418)
                        (void)core::plane_de(total, dz, FE{p.spacing_m, p.spacing_e}, dx, dy);
                    }
                }
                planes.de_x[index] = dx;
                planes.de_y[index] = dy;
                planes.test1[index] = test1;
                planes.test2[index] = test2;
            }
            return;
        }
    }
    p.output[index] = core::encode_iteration(p.max_iter, p.output_bias);
    if constexpr (Planes) core::clear_planes(planes, index, p.max_iter);
}

}  // namespace direct
}  // namespace fav
