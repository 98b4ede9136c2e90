// kfp_core.h -- per-pixel Kalles SetColor transfer of a scalar iteration
// field and, when given, its Kalles planes: orbit / stored iteration,
// bailout, transition, phase, analytic DE, |z|^2 tests and a texture
// (renderer.cpp:5655-5976, 6696-7289).
//
// Compiled for the gfx950 colour kernel (kfp_kernel.hip) and, unchanged, for
// the test-only CPU harness.  Double and float stages follow the reference's
// types and operation order (-ffp-contract=off); the transcendental
// functions are the platform's, which is why device output may differ from
// the CPU by one 8-bit level at a quantisation boundary.
#pragma once

#include <cstdint>

#include "fav_numeric.h"

namespace fav {
namespace kfp {

constexpr double kPi = 3.14159265358979323846;

// Per-launch constants (kfp_pixel_context / kfp_slope_direction /
// kfp_transfer_bounds results, renderer.cpp:5723-5763, 7296-7334, 7509-7514).
struct KfpParams {
    const float* field;
    const uint8_t* lut;
    uint8_t* output;
    const double* bounds;  // [0] minimum, [1] maximum for colour method 4
    int32_t width, height, max_iter, lut_size;
    double field_bias, smooth_offset, iter_div, color_offset;
    double slope_power, slope_ratio, cosine, sine;
    int32_t color_method, differences, smooth, flat, inverse_transition;
    int32_t needs_difference, needs_slopes;
    int32_t multi_color, blend_multi_color, multi_color_count;
    int32_t interior_color[3];
    double multi_color_period[256];
    int32_t multi_color_type[256];
    // FractalKfpPlanes (each null when absent; has_planes when a planes
    // struct was given at all) and the kfp_pixel_context constants
    // (renderer.cpp:5655-5763) evaluated on the host
    int32_t has_planes;
    const int64_t* orbit_iteration;
    const int64_t* iteration;
    const double* bailout;
    const double* transition;
    const double* phase;
    const double* de_x;
    const double* de_y;
    const double* test1;
    const double* test2;
    const uint8_t* texture;
    int32_t texture_width, texture_height, texture_stride;
    int32_t smooth_method, has_constant_bailout, has_smooth_bailout_log, has_smooth_power_log;
    int32_t has_smooth_norm_exponent, texture_active, show_glitches;
    double power, selected_bailout_radius, smooth_bailout_radius, smooth_bailout_log, smooth_power_log;
    double smooth_norm_exponent, phase_color_strength, texture_merge, texture_power, texture_ratio;
};

FAV_HD double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
FAV_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
FAV_HD double maxd(double a, double b) { return a < b ? b : a; }  // std::max
FAV_HD double mind(double a, double b) { return b < a ? b : a; }  // std::min

// kfp_colour_dither_mask + kfp_dithered_srgb_byte, renderer.cpp:5246-5336.
FAV_HD uint8_t dithered_srgb_byte(float value, int x, int y, int channel) {
    if (!__builtin_isfinite(value)) value = 0.0F;
    const uint32_t coordinate = (static_cast<uint32_t>(x) + static_cast<uint32_t>(channel) * 67U
                                 + static_cast<uint32_t>(y) * 236U) * 119U;
    const float mask = static_cast<float>(coordinate & 255U) / 256.0F;
    const float quantized = 255.0F * value + mask;
    return static_cast<uint8_t>(clampi(static_cast<int>(::floorf(quantized)), 0, 255));
}

FAV_HD bool inside_value(float value, int max_iter, double bias) {  // renderer.cpp:5379-5387
    return __builtin_isfinite(value) && static_cast<double>(value) >= static_cast<double>(max_iter) - bias;
}

// kfp_colour_sample, renderer.cpp:5434-5460 (the materialised stencil value).
FAV_HD double colour_sample(const KfpParams& k, int x, int y) {
    x = clampi(x, 0, k.width - 1);
    y = clampi(y, 0, k.height - 1);
    const float value = k.field[static_cast<size_t>(y) * static_cast<size_t>(k.width) + static_cast<size_t>(x)];
    if (!__builtin_isfinite(value)) return maxd(0.0, k.smooth_offset);
    if (static_cast<double>(value) >= static_cast<double>(k.max_iter) - k.field_bias)
        return static_cast<double>(k.max_iter) + 1.0;
    return maxd(0.0, clampd(static_cast<double>(value) + k.field_bias, 0.0, static_cast<double>(k.max_iter))
                         + k.smooth_offset);
}

FAV_HD double reflected(const KfpParams& k, int x, int y, int ox, int oy, double centre) {
    const int sx = x + ox, sy = y + oy;
    if (sx >= 0 && sx < k.width && sy >= 0 && sy < k.height) return colour_sample(k, sx, sy);
    const int px = x - ox, py = y - oy;
    if (px >= 0 && px < k.width && py >= 0 && py < k.height) return 2.0 * centre - colour_sample(k, px, py);
    return centre;
}

FAV_HD bool has_iteration_planes(const KfpParams& k) { return k.orbit_iteration || k.iteration; }

// kfp_plane_bailout, renderer.cpp:5765-5774.
FAV_HD double plane_bailout(const KfpParams& k, size_t i) {
    if (k.bailout) {
        const double v = k.bailout[i];
        if (__builtin_isfinite(v) && v > 1.0) return v;
    }
    return k.selected_bailout_radius;
}

// kfp_plane_smooth_part with the pixel context, renderer.cpp:5776-5837.
FAV_HD double plane_smooth_part(const KfpParams& k, size_t i) {
    if (!k.test1) return 0.0;
    const double test1 = k.test1[i];
    if (!__builtin_isfinite(test1) || test1 <= 0.0) return 0.0;
    double smooth = 0.0;
    if (k.smooth_method == 0) {
        const double radius = k.has_constant_bailout ? k.smooth_bailout_radius : plane_bailout(k, i);
        if (radius > 1.0 && __builtin_isfinite(k.power) && k.power > 1.0) {
            const double numerator = ::log(::sqrt(test1));
            const double denominator = k.has_smooth_bailout_log ? k.smooth_bailout_log : ::log(radius);
            const double power_log = k.has_smooth_power_log ? k.smooth_power_log : ::log(k.power);
            if (__builtin_isfinite(numerator) && numerator > 0.0 && __builtin_isfinite(denominator)
                && denominator != 0.0 && __builtin_isfinite(power_log) && power_log != 0.0)
                smooth = 1.0 - ::log(numerator / denominator) / power_log;
        }
    } else if (k.smooth_method == 1 && k.test2) {
        const double test2 = k.test2[i];
        if (__builtin_isfinite(test2) && test2 >= 0.0 && k.has_smooth_norm_exponent) {
            const double root1 = ::pow(test1, k.smooth_norm_exponent);
            const double root2 = ::pow(test2, k.smooth_norm_exponent);
            const double denominator = root1 - root2;
            const double radius = k.has_constant_bailout ? k.smooth_bailout_radius : plane_bailout(k, i);
            if (__builtin_isfinite(root1) && __builtin_isfinite(root2) && __builtin_isfinite(radius)
                && denominator != 0.0)
                smooth = 1.0 - (root1 - radius) / denominator;
        }
    }
    return __builtin_isfinite(smooth) ? smooth : 0.0;
}

// kfp_plane_orbit_sample, renderer.cpp:5844-5864: the continuous colour
// sample of an orbit-plane pixel, max_iter + 1 for interior.
FAV_HD double orbit_sample(const KfpParams& k, size_t i, bool& inside) {
    const int64_t raw = k.orbit_iteration[i] > 0 ? k.orbit_iteration[i] : 0;
    if (raw >= k.max_iter) {
        inside = true;
        return static_cast<double>(k.max_iter) + 1.0;
    }
    const double combined = static_cast<double>(raw) + plane_smooth_part(k, i);
    inside = combined >= static_cast<double>(k.max_iter);
    return inside ? static_cast<double>(k.max_iter) + 1.0 : combined;
}

// kfp_plane_colour_sample, renderer.cpp:5866-5905 (with the context's
// smooth offset).
FAV_HD double plane_colour_sample(const KfpParams& k, int x, int y) {
    x = clampi(x, 0, k.width - 1);
    y = clampi(y, 0, k.height - 1);
    const size_t i = static_cast<size_t>(y) * static_cast<size_t>(k.width) + static_cast<size_t>(x);
    if (k.orbit_iteration) {
        bool inside;
        return orbit_sample(k, i, inside);
    }
    if (k.iteration) {
        const int64_t stored = k.iteration[i];
        if (stored >= k.max_iter) return static_cast<double>(k.max_iter) + 1.0;
        const double t = k.transition ? k.transition[i] : 1.0;
        return static_cast<double>(stored > 0 ? stored : 0) + 1.0 - (__builtin_isfinite(t) ? t : 1.0);
    }
    return colour_sample(k, x, y);
}

// The stencil sample of write_kfp_pixel's `reflected` (renderer.cpp:6834-6889).
FAV_HD double reflected_sample(const KfpParams& k, int x, int y, int ox, int oy, double centre) {
    if (!has_iteration_planes(k)) return reflected(k, x, y, ox, oy, centre);
    const int sx = x + ox, sy = y + oy;
    if (sx >= 0 && sx < k.width && sy >= 0 && sy < k.height) return plane_colour_sample(k, sx, sy);
    const int px = x - ox, py = y - oy;
    if (px >= 0 && px < k.width && py >= 0 && py < k.height) return 2.0 * centre - plane_colour_sample(k, px, py);
    return centre;
}

FAV_HD double plane_value(const double* v, size_t i) {  // kfp_plane_value, renderer.cpp:5936-5943
    if (!v) return 0.0;
    return __builtin_isfinite(v[i]) ? v[i] : 0.0;
}

// kfp_texture_index, renderer.cpp:5971-5976.
FAV_HD int texture_index(double value, int limit) {
    if (limit <= 0 || !__builtin_isfinite(value)) return 0;
    if (value <= 0.0) return 0;
    if (value >= static_cast<double>(limit - 1)) return limit - 1;
    return static_cast<int>(value);
}

// kfp_difference_magnitude, renderer.cpp:5497-5589.
FAV_HD double difference_magnitude(int d, double c, double l, double r, double u, double dn, double tl, double tr,
                                   double bl, double br) {
    const double inverse_sqrt_two = 0.7071067811865475244008443621048490;
    if (d == 0) return __builtin_fabs(l - c) * 1.414 + __builtin_fabs(u - c) * 1.414 + __builtin_fabs(tl - c)
        + __builtin_fabs(bl - c);
    if (d == 1) {
        const double sq = (l - c) * (l - c) + (r - c) * (r - c) + (u - c) * (u - c) + (dn - c) * (dn - c)
            + ((tl - c) * (tl - c) + (br - c) * (br - c)) * inverse_sqrt_two * inverse_sqrt_two
            + ((bl - c) * (bl - c) + (tr - c) * (tr - c)) * inverse_sqrt_two * inverse_sqrt_two;
        return ::sqrt(maxd(0.0, sq * 0.25)) * 2.8284271247461903;
    }
    if (d == 2) {
        const double sq = (r - l) * (r - l) * 0.25 + (dn - u) * (dn - u) * 0.25 + (br - tl) * (br - tl) * 0.125
            + (tr - bl) * (tr - bl) * 0.125;
        return ::sqrt(maxd(0.0, sq * 0.5)) * 2.8284271247461903;
    }
    if (d == 3) {
        const double sq = (tl - c) * (tl - c) / 2.0 + (l - u) * (l - u) / 2.0;
        return ::sqrt(maxd(0.0, sq)) * 2.8284271247461903;
    }
    if (d == 4) {
        const double dx = ((u - tl) + (c - l)) * 0.5;
        const double dy = ((l - tl) + (c - u)) * 0.5;
        return ::sqrt(dx * dx + dy * dy) * 2.8284271247461903;
    }
    if (d == 5) {
        const double dx = (r + tr + br - l - tl - bl) / 6.0;
        const double dy = (dn + bl + br - u - tl - tr) / 6.0;
        return ::sqrt(dx * dx + dy * dy) * 2.8284271247461903;
    }
    if (d == 6) {
        const double lap = tl + 4.0 * u + tr + 4.0 * l - 20.0 * c + 4.0 * r + bl + 4.0 * dn + br;
        return ::sqrt(__builtin_fabs(lap / 6.0 * 1.4426950408889634)) * 2.8284271247461903;
    }
    return 0.0;  // Differences_Analytic without DE planes
}

// kfp_hsv_to_rgb, renderer.cpp:5591-5644 (float, as Kalles' colour.h).
FAV_HD void hsv_to_rgb(float hue, float saturation, float value, float& r, float& g, float& b) {
    const float scaled_hue = hue * 6.0F;
    const int sector = static_cast<int>(::floorf(scaled_hue));
    float fraction = scaled_hue - static_cast<float>(sector);
    if ((sector & 1) == 0) fraction = 1.0F - fraction;
    const float minimum = value * (1.0F - saturation);
    const float transition = value * (1.0F - saturation * fraction);
    r = 0.0F; g = 0.0F; b = 0.0F;
    switch (sector) {
        case 0: case 6: r = minimum; g = transition; b = value; break;
        case 1: r = minimum; g = value; b = transition; break;
        case 2: r = transition; g = value; b = minimum; break;
        case 3: r = value; g = transition; b = minimum; break;
        case 4: r = value; g = minimum; b = transition; break;
        case 5: r = transition; g = minimum; b = value; break;
        default: break;
    }
}

// palette_rgb, renderer.cpp:7070-7110: cyclic LUT lookup with optional
// linear blend of the two neighbouring entries.
FAV_HD void palette_rgb(const KfpParams& k, double input, float& r, float& g, float& b) {
    double position = ::fmod(input, static_cast<double>(k.lut_size));
    if (position < 0.0) position += static_cast<double>(k.lut_size);
    const double lower_position = ::floor(position);
    const int lower = clampi(static_cast<int>(lower_position), 0, k.lut_size - 1);
    const uint8_t* lut = k.lut;
    if (k.smooth) {
        double fraction = position - lower_position;
        if (k.inverse_transition) fraction = 1.0 - fraction;
        const int upper = (lower + 1) % k.lut_size;
        const double inv = 1.0 - fraction;
        r = static_cast<float>((static_cast<double>(lut[lower * 3]) * inv + static_cast<double>(lut[upper * 3]) * fraction)
                               / 255.0F);
        g = static_cast<float>((static_cast<double>(lut[lower * 3 + 1]) * inv
                                + static_cast<double>(lut[upper * 3 + 1]) * fraction) / 255.0F);
        b = static_cast<float>((static_cast<double>(lut[lower * 3 + 2]) * inv
                                + static_cast<double>(lut[upper * 3 + 2]) * fraction) / 255.0F);
    } else {
        r = static_cast<float>(lut[lower * 3]) / 255.0F;
        g = static_cast<float>(lut[lower * 3 + 1]) / 255.0F;
        b = static_cast<float>(lut[lower * 3 + 2]) / 255.0F;
    }
}

// write_kfp_pixel, renderer.cpp:6696-7289: the pixel (x, y) of the field,
// dithered at the output position (dither_x, dither_y), with distances and
// slopes scaled by the output width `spatial_width`.
FAV_HD void colour_at(const KfpParams& k, int x, int y, int dither_x, int dither_y, int spatial_width, uint8_t* dst) {
    const size_t idx = static_cast<size_t>(y) * static_cast<size_t>(k.width) + static_cast<size_t>(x);
    if (k.transition && k.transition[idx] < 0.0 && !k.show_glitches) {  // a hidden glitch pixel
        dst[0] = 0;
        dst[1] = 0;
        dst[2] = 0;
        return;
    }
    bool inside = false;
    double smooth_iter;
    if (k.orbit_iteration) {
        smooth_iter = orbit_sample(k, idx, inside);
    } else {
        const float raw = k.field[idx];
        if (k.iteration) {
            inside = k.iteration[idx] >= k.max_iter;  // kfp_plane_inside, renderer.cpp:5907-5934
            smooth_iter = plane_colour_sample(k, x, y);
        } else {
            inside = inside_value(raw, k.max_iter, k.field_bias);
            const double safe = __builtin_isfinite(raw)
                ? clampd(static_cast<double>(raw) + k.field_bias, 0.0, static_cast<double>(k.max_iter)) : 0.0;
            smooth_iter = maxd(0.0, safe + k.smooth_offset);
        }
    }
    if (inside && !k.texture_active) {
        dst[0] = static_cast<uint8_t>(k.interior_color[0]);
        dst[1] = static_cast<uint8_t>(k.interior_color[1]);
        dst[2] = static_cast<uint8_t>(k.interior_color[2]);
        return;
    }
    const double colour_iter = k.flat ? ::floor(smooth_iter) : smooth_iter;
    const double c = smooth_iter;
    double l = c, r = c, u = c, dn = c, tl = c, tr = c, bl = c, br = c;
    if (k.needs_difference || k.needs_slopes || k.texture_active) {
        l = reflected_sample(k, x, y, -1, 0, c);
        u = reflected_sample(k, x, y, 0, -1, c);
        if (k.needs_difference) {
            switch (k.differences) {
                case 0:
                    tl = reflected_sample(k, x, y, -1, -1, c);
                    bl = reflected_sample(k, x, y, -1, 1, c);
                    break;
                case 1: case 2: case 5: case 6:
                    r = reflected_sample(k, x, y, 1, 0, c);
                    dn = reflected_sample(k, x, y, 0, 1, c);
                    tl = reflected_sample(k, x, y, -1, -1, c);
                    tr = reflected_sample(k, x, y, 1, -1, c);
                    bl = reflected_sample(k, x, y, -1, 1, c);
                    br = reflected_sample(k, x, y, 1, 1, c);
                    break;
                case 3: case 4:
                    tl = reflected_sample(k, x, y, -1, -1, c);
                    break;
                default:
                    r = reflected_sample(k, x, y, 1, 0, c);
                    dn = reflected_sample(k, x, y, 0, 1, c);
                    break;
            }
        }
        if (k.needs_slopes || k.texture_active) {
            if (x == 0) r = reflected_sample(k, x, y, 1, 0, c);
            if (y == 0) dn = reflected_sample(k, x, y, 0, 1, c);
        }
    }
    // analytic slopes from the DE planes in float (renderer.cpp:6943-6962)
    const bool analytic = k.differences == 7 && k.de_x && k.de_y;
    double slope_x = 0.0, slope_y = 0.0, gradient = 0.0;
    if ((k.needs_difference || k.needs_slopes) && analytic) {
        const float dx = static_cast<float>(plane_value(k.de_x, idx));
        const float dy = static_cast<float>(plane_value(k.de_y, idx));
        const float denominator = dx * dx + dy * dy;
        if (denominator > 0.0F && __builtin_isfinite(denominator)) {
            gradient = static_cast<double>(1.0F / ::sqrtf(denominator));
            slope_x = static_cast<double>(dx / denominator);
            slope_y = static_cast<double>(-dy / denominator);
        }
    } else if (k.needs_difference) {
        gradient = difference_magnitude(k.differences, c, l, r, u, dn, tl, tr, bl, br);
    }
    double distance = __builtin_isnan(gradient) ? 0.0 : gradient * static_cast<double>(spatial_width) / 640.0;
    if (distance < 0.0) distance = 0.0;
    double transfer = colour_iter;
    switch (k.color_method) {
        case 1: transfer = ::sqrt(maxd(0.0, colour_iter)); break;
        case 2: transfer = ::pow(maxd(0.0, colour_iter), 1.0 / 3.0); break;
        case 3: transfer = ::log(maxd(1.0, colour_iter)); break;
        case 4:
            transfer = 1024.0 * (colour_iter - k.bounds[0]) / maxd(k.bounds[1] - k.bounds[0], 1.0e-12);
            break;
        case 5: transfer = mind(distance, 1024.0); break;
        case 6: {
            const double dt = mind(::sqrt(maxd(0.0, distance)), 1024.0);
            transfer = dt > k.iter_div ? smooth_iter : dt;
            break;
        }
        case 7: transfer = ::log(maxd(1.0, distance + 1.0)); break;
        case 8: transfer = ::sqrt(maxd(0.0, distance)); break;
        case 9: transfer = ::log(1.0 + ::log(1.0 + maxd(0.0, colour_iter))); break;
        case 10: transfer = ::atan(colour_iter); break;
        case 11: transfer = ::sqrt(::sqrt(maxd(0.0, colour_iter))); break;
        default: break;
    }
    if (__builtin_isnan(transfer) || transfer < 0.0) transfer = 0.0;
    if (k.color_method == 5 || k.color_method == 7 || k.color_method == 8) transfer = clampd(transfer, 0.0, 1024.0);
    const double phase_shift = k.phase ? k.phase_color_strength / 100.0 * 1024.0 * plane_value(k.phase, idx) : 0.0;
    double palette_value = transfer;
    if (k.iter_div != 1.0) palette_value /= k.iter_div;
    const double base = palette_value + k.color_offset;
    float red, green, blue;
    if (k.multi_color) {
        const double wave_input = k.smooth ? base : ::floor(base);
        double hs = 0.0, ss = 0.0, vs = 0.0;
        int hc = 0, sc = 0, vc = 0;
        for (int i = 0; i < k.multi_color_count; ++i) {
            const double period = k.multi_color_period[i];
            const double wave = period < 0.0 ? -period / 100.0 : 0.5 + 0.5 * ::sin(kPi * wave_input / period);
            if (k.multi_color_type[i] == 0) { hs += wave; ++hc; }
            else if (k.multi_color_type[i] == 1) { ss += wave; ++sc; }
            else { vs += wave; ++vc; }
        }
        float mr, mg, mb;
        hsv_to_rgb(static_cast<float>(hc > 0 ? hs / hc : 0.0), static_cast<float>(sc > 0 ? ss / sc : 0.0),
                   static_cast<float>(vc > 0 ? vs / vc : 0.0), mr, mg, mb);
        if (k.blend_multi_color) {
            palette_rgb(k, base + phase_shift, red, green, blue);
            red = (red + mr) * 0.5F;
            green = (green + mg) * 0.5F;
            blue = (blue + mb) * 0.5F;
        } else {
            red = mr; green = mg; blue = mb;
        }
    } else {
        palette_rgb(k, base + phase_shift, red, green, blue);
    }
    if (k.texture_active && inside) {  // textured interior starts from the interior colour
        red = static_cast<float>(k.interior_color[0]) / 255.0F;
        green = static_cast<float>(k.interior_color[1]) / 255.0F;
        blue = static_cast<float>(k.interior_color[2]) / 255.0F;
    }
    if (k.texture_active) {  // CFraktalSFT::SetTexture, renderer.cpp:7190-7252
        const double tdx = x > 0 ? l - c : (x + 1 < k.width ? c - r : 0.0);
        const double tdy = y > 0 ? u - c : (y + 1 < k.height ? c - dn : 0.0);
        auto offset = [&](double difference) {
            const double powered = ::pow(1.0 + difference, k.texture_power);
            const bool forward = powered > 1.0;
            const double selected = forward ? powered : 1.0 / powered;
            const double mapped = !__builtin_isnan(selected)
                ? (::atan(selected) - kPi / 4.0) / (kPi / 4.0) * (k.texture_ratio / 100.0) : 0.0;
            return __builtin_isfinite(mapped) ? (forward ? k.texture_power * mapped : -k.texture_power * mapped)
                                              : 0.0;
        };
        int image_offset = 0;
        if (__builtin_isfinite(k.texture_power))
            image_offset = static_cast<int>(clampd(k.texture_power / 64.0, -2147483648.0, 2147483647.0));
        const int sx = texture_index(static_cast<double>(x + image_offset) + offset(tdx), k.texture_width);
        const int sy = texture_index(static_cast<double>(y + image_offset) + offset(tdy), k.texture_height);
        const int row = k.texture_height - 1 - sy;
        const uint8_t* t = k.texture + static_cast<size_t>(row) * static_cast<size_t>(k.texture_stride)
            + static_cast<size_t>(sx) * 3U;
        const double m = k.texture_merge;
        red = static_cast<float>(static_cast<double>(red) * (1.0 - m) + m * static_cast<double>(t[0]) / 255.0);
        green = static_cast<float>(static_cast<double>(green) * (1.0 - m) + m * static_cast<double>(t[1]) / 255.0);
        blue = static_cast<float>(static_cast<double>(blue) * (1.0 - m) + m * static_cast<double>(t[2]) / 255.0);
    }
    if (k.needs_slopes) {
        const double horizontal = analytic ? slope_x : (x > 0 ? l - c : c - r);
        const double vertical = analytic ? slope_y : (y == 0 ? c - dn : u - c);
        const double projected = ((horizontal * k.cosine) + (vertical * k.sine)) * k.slope_power
            * static_cast<double>(spatial_width) / 640.0;
        const double strength = ::atan(__builtin_fabs(projected)) / (kPi / 2.0) * k.slope_ratio / 100.0;
        if (projected >= 0.0) {
            red = static_cast<float>(static_cast<double>(red) * (1.0 - strength));
            green = static_cast<float>(static_cast<double>(green) * (1.0 - strength));
            blue = static_cast<float>(static_cast<double>(blue) * (1.0 - strength));
        } else {
            red = static_cast<float>(static_cast<double>(red) * (1.0 - strength) + strength);
            green = static_cast<float>(static_cast<double>(green) * (1.0 - strength) + strength);
            blue = static_cast<float>(static_cast<double>(blue) * (1.0 - strength) + strength);
        }
    }
    dst[0] = dithered_srgb_byte(red, dither_x, dither_y, 0);
    dst[1] = dithered_srgb_byte(green, dither_x, dither_y, 1);
    dst[2] = dithered_srgb_byte(blue, dither_x, dither_y, 2);
}

// A frame pixel: field, dither and output at the same position.
FAV_HD void colour_pixel(const KfpParams& k, int x, int y) {
    const size_t idx = static_cast<size_t>(y) * static_cast<size_t>(k.width) + static_cast<size_t>(x);
    colour_at(k, x, y, x, y, k.width, k.output + idx * 3U);
}

// fractal_atlas_colourise_kfp's geometry (renderer.cpp:16750-17321): an
// output-sized parent and a child tile at (child_left, child_top), each
// coloured from its own field and transfer range.
struct RectAtlas {
    int32_t use_child, child_left, child_top, feather;
    const float* edge_x;  // separable seam weights of the child columns / rows
    const float* edge_y;
};

FAV_HD uint8_t rounded_byte(double v) {  // rounded_colour_byte, renderer.cpp:5237-5242
    v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
    const int i = static_cast<int>(__builtin_rint(v));
    return static_cast<uint8_t>(i < 0 ? 0 : (i > 255 ? 255 : i));
}

// One output pixel of fractal_atlas_colourise_kfp (renderer.cpp:17080-17305).
FAV_HD void rect_atlas_pixel(const KfpParams& parent, const KfpParams& child, const RectAtlas& g, int x, int y) {
    uint8_t* dst = parent.output + (static_cast<size_t>(y) * static_cast<size_t>(parent.width) + static_cast<size_t>(x)) * 3U;
    const int ow = parent.width;
    if (!g.use_child || x < g.child_left || x >= g.child_left + child.width || y < g.child_top
        || y >= g.child_top + child.height) {
        colour_at(parent, x, y, x, y, ow, dst);
        return;
    }
    const int cx = x - g.child_left, cy = y - g.child_top;
    const float child_raw = child.field[static_cast<size_t>(cy) * static_cast<size_t>(child.width) + static_cast<size_t>(cx)];
    const bool child_inside = inside_value(child_raw, child.max_iter, child.field_bias);
    const float alpha = g.feather >= 2 ? (g.edge_x[cx] < g.edge_y[cy] ? g.edge_x[cx] : g.edge_y[cy]) : 1.0F;
    if (child_inside && alpha >= 0.999999F) {
        colour_at(child, cx, cy, x, y, ow, dst);
        return;
    }
    if (g.feather < 2) {
        const float parent_raw = parent.field[static_cast<size_t>(y) * static_cast<size_t>(ow) + static_cast<size_t>(x)];
        if (!inside_value(parent_raw, parent.max_iter, parent.field_bias)) colour_at(parent, x, y, x, y, ow, dst);
        else colour_at(child, cx, cy, x, y, ow, dst);
        return;
    }
    if (alpha >= 0.999999F) {
        colour_at(child, cx, cy, x, y, ow, dst);
        return;
    }
    uint8_t p[3], c[3];
    colour_at(parent, x, y, x, y, ow, p);
    colour_at(child, cx, cy, x, y, ow, c);
    const double a = static_cast<double>(alpha);
    for (int ch = 0; ch < 3; ++ch)
        dst[ch] = rounded_byte(static_cast<double>(p[ch]) * (1.0 - a) + static_cast<double>(c[ch]) * a);
}

// The value whose range kfp_transfer_bounds takes (renderer.cpp:7296-7334):
// false when the pixel is excluded.  The value is an integer in
// [floor(smooth_offset), max_iter - 2].
FAV_HD bool transfer_bound_value(const KfpParams& k, size_t index, long long& value) {
    if (k.has_planes) {  // kfp_transfer_bounds_planes, renderer.cpp:7336-7394
        double v;
        if (k.orbit_iteration) {
            bool inside;
            const double colour = orbit_sample(k, index, inside);
            if (colour > static_cast<double>(k.max_iter)) return false;
            v = ::floor(maxd(0.0, colour));
        } else {
            const int y = static_cast<int>(index / static_cast<size_t>(k.width));
            const int x = static_cast<int>(index - static_cast<size_t>(y) * static_cast<size_t>(k.width));
            if (k.iteration ? k.iteration[index] >= k.max_iter : inside_value(k.field[index], k.max_iter, k.field_bias))
                return false;
            const double colour = plane_colour_sample(k, x, y);  // kfp_plane_stored_iteration
            v = colour > static_cast<double>(k.max_iter) ? static_cast<double>(k.max_iter) : ::floor(maxd(0.0, colour));
        }
        if (v >= static_cast<double>(k.max_iter - 1)) return false;
        value = static_cast<long long>(v);
        return true;
    }
    const float raw = k.field[index];
    if (inside_value(raw, k.max_iter, k.field_bias)) return false;
    const double safe = __builtin_isfinite(raw)
        ? clampd(static_cast<double>(raw) + k.field_bias, 0.0, static_cast<double>(k.max_iter)) : 0.0;
    const double v = ::floor(safe + k.smooth_offset);
    if (v >= static_cast<double>(k.max_iter - 1)) return false;
    value = static_cast<long long>(v);
    return true;
}

}  // namespace kfp
}  // namespace fav
