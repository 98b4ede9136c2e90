// kfp_prepare.h -- host-side validation of FractalKfpOptions and the
// per-launch constants of the KFP colour pass (shared by the C ABI and the
// CPU harness of the kernel code).
#pragma once

#include <cmath>
#include <cstddef>
#include <stdexcept>

#include "../../include/fractal_hip.h"
#include "kfp_core.h"

namespace fav {
namespace kfp {

// valid_kfp_options, renderer.cpp:7396-7454.
inline bool valid_options(const FractalKfpOptions* o, int lut_size) {
    const uint32_t legacy_size = static_cast<uint32_t>(offsetof(FractalKfpOptions, show_glitches));
    if (!o || o->struct_size < legacy_size || o->version != FRACTAL_KFP_OPTIONS_VERSION || lut_size < 2
        || lut_size > 65536 || !std::isfinite(o->iter_div) || o->iter_div <= 0.0 || !std::isfinite(o->color_offset)
        || !std::isfinite(o->ratio) || o->color_method < 0 || o->color_method > 11 || o->smooth_method < 0
        || o->smooth_method > 2 || o->smooth < 0 || o->smooth > 1 || o->flat < 0 || o->flat > 1
        || o->inverse_transition < 0 || o->inverse_transition > 1 || !std::isfinite(o->phase_color_strength)
        || o->multi_color < 0 || o->multi_color > 1 || o->blend_multi_color < 0 || o->blend_multi_color > 1
        || o->multi_color_count > FRACTAL_KFP_MAX_MULTI_COLORS || !std::isfinite(o->power) || o->power <= 0.0
        || o->slopes < 0 || o->slopes > 1 || !std::isfinite(o->slope_power) || o->slope_power < 0.0
        || !std::isfinite(o->slope_ratio) || o->slope_ratio < 0.0 || !std::isfinite(o->slope_angle)
        || o->differences < 0 || o->differences > 7 || !std::isfinite(o->field_bias) || o->field_bias < 0.0
        || o->bailout_radius_preset < 0 || o->bailout_radius_preset > 3 || !std::isfinite(o->bailout_radius_custom)
        || o->bailout_radius_custom <= 0.0 || o->bailout_norm_preset < 0 || o->bailout_norm_preset > 3
        || !std::isfinite(o->bailout_norm_custom) || o->bailout_norm_custom <= 0.0 || o->texture_enabled < 0
        || o->texture_enabled > 1 || !std::isfinite(o->texture_merge) || !std::isfinite(o->texture_power)
        || !std::isfinite(o->texture_ratio) || o->texture_resize < 0 || o->texture_resize > 1 || o->use_opengl < 0
        || o->use_opengl > 1 || o->use_srgb < 0 || o->use_srgb > 1)
        return false;
    if (o->struct_size >= legacy_size + sizeof(o->show_glitches) && (o->show_glitches < 0 || o->show_glitches > 1))
        return false;
    for (uint32_t i = 0; i < o->multi_color_count; ++i)
        if (!std::isfinite(o->multi_color_period[i]) || o->multi_color_period[i] == 0.0
            || o->multi_color_type[i] < 0 || o->multi_color_type[i] > 2)
            return false;
    return o->interior_color[0] >= 0 && o->interior_color[0] <= 255 && o->interior_color[1] >= 0
        && o->interior_color[1] <= 255 && o->interior_color[2] >= 0 && o->interior_color[2] <= 255;
}

// kfp_smooth_offset, renderer.cpp:5400-5428 (evaluated on the host, once).
inline double smooth_offset(const FractalKfpOptions& o) {
    if (o.smooth_method != 0 || !std::isfinite(o.power) || o.power <= 0.0 || std::abs(o.power - 1.0) < 1.0e-12)
        return 0.0;
    double radius = 10000.0;
    switch (o.bailout_radius_preset) {
        case 0: radius = 10000.0; break;
        case 1: radius = 2.0; break;
        case 2: radius = std::pow(2.0, 1.0 / (o.power - 1.0)); break;
        case 3: radius = o.bailout_radius_custom; break;
        default: return 0.0;
    }
    if (!std::isfinite(radius) || radius <= 1.0) return 0.0;
    return std::log(std::log(radius)) / std::log(o.power);
}

// Fills the launch constants from validated options (the per-image parts of
// fractal_colourise_kfp_impl, renderer.cpp:16314-16545).  Throws when the
// field bias exceeds the iteration cap, as the reference does.
inline KfpParams make_params(const FractalKfpOptions& o, const float* field, uint8_t* output, int width, int height,
                             int max_iter, const uint8_t* lut, int lut_size) {
    if (o.field_bias > static_cast<double>(max_iter))
        throw std::runtime_error("native KFP field bias exceeds iteration cap");
    KfpParams k{};
    k.field = field;
    k.lut = lut;
    k.output = output;
    k.width = width;
    k.height = height;
    k.max_iter = max_iter;
    k.lut_size = lut_size;
    k.field_bias = o.field_bias;
    k.smooth_offset = smooth_offset(o);
    k.iter_div = o.iter_div;
    k.color_offset = o.color_offset;
    k.slope_power = o.slope_power;
    k.slope_ratio = o.slope_ratio;
    k.cosine = 1.0;
    k.sine = 0.0;
    if (o.slopes) {  // kfp_slope_direction, renderer.cpp:7509-7514
        const double angle = o.slope_angle * kPi / 180.0;
        k.cosine = std::cos(angle);
        k.sine = std::sin(angle);
    }
    k.color_method = o.color_method;
    k.differences = o.differences;
    k.smooth = o.smooth;
    k.flat = o.flat;
    k.inverse_transition = o.inverse_transition;
    k.needs_difference = o.color_method >= 5 && o.color_method <= 8;
    k.needs_slopes = o.slopes && o.slope_power > 0.0 && o.slope_ratio > 0.0;
    k.multi_color = o.multi_color;
    k.blend_multi_color = o.blend_multi_color;
    k.multi_color_count = static_cast<int32_t>(o.multi_color_count);
    for (int i = 0; i < 3; ++i) k.interior_color[i] = o.interior_color[i];
    for (uint32_t i = 0; i < o.multi_color_count; ++i) {
        k.multi_color_period[i] = o.multi_color_period[i];
        k.multi_color_type[i] = o.multi_color_type[i];
    }
    return k;
}

// kfp_selected_bailout_radius / _norm, renderer.cpp:5672-5716.
inline double selected_bailout_radius(const FractalKfpOptions& o) {
    double radius = 10000.0;
    switch (o.bailout_radius_preset) {
        case 0: radius = 10000.0; break;
        case 1: radius = 2.0; break;
        case 2:
            if (o.power > 1.0 && std::isfinite(o.power)) radius = std::pow(2.0, 1.0 / (o.power - 1.0));
            else return 0.0;
            break;
        case 3: radius = o.bailout_radius_custom; break;
        default: return 0.0;
    }
    return std::isfinite(radius) && radius > 1.0 ? radius : 0.0;
}
inline double selected_bailout_norm(const FractalKfpOptions& o) {
    switch (o.bailout_norm_preset) {
        case 0: return 1.0;
        case 1: return 2.0;
        case 2: return HUGE_VAL;
        case 3: return std::isfinite(o.bailout_norm_custom) && o.bailout_norm_custom > 0.0 ? o.bailout_norm_custom : 0.0;
        default: return 0.0;
    }
}

// valid_kfp_planes, renderer.cpp:7467-7491.
inline bool valid_planes(const FractalKfpPlanes* p, int width, int height) {
    if (!p) return true;
    if (p->struct_size < sizeof(FractalKfpPlanes) || p->version != FRACTAL_KFP_PLANES_VERSION) return false;
    if (!p->texture_rgb) return p->texture_width == 0 && p->texture_height == 0 && p->texture_stride == 0;
    if (p->texture_width <= 0 || p->texture_height <= 0
        || static_cast<int64_t>(p->texture_stride) < static_cast<int64_t>(p->texture_width) * 3)
        return false;
    const uint64_t bytes = static_cast<uint64_t>(p->texture_stride) * static_cast<uint64_t>(p->texture_height);
    return bytes <= 100000000ULL * 3U && width > 0 && height > 0
        && static_cast<uint64_t>(width) <= 100000000ULL / static_cast<uint64_t>(height);
}

// The plane-dependent launch constants: kfp_pixel_context
// (renderer.cpp:5723-5763) and kfp_show_glitches (7456-7465).  Plane
// pointers are left null: the caller binds the device copies.
inline void set_plane_context(KfpParams& k, const FractalKfpOptions& o, const FractalKfpPlanes* planes) {
    k.has_planes = planes != nullptr;
    k.smooth_method = o.smooth_method;
    k.power = o.power;
    k.has_constant_bailout = planes != nullptr && planes->bailout == nullptr;
    k.selected_bailout_radius = selected_bailout_radius(o);
    k.smooth_bailout_radius = k.selected_bailout_radius;
    if (std::isfinite(o.power) && o.power > 1.0) {
        k.smooth_power_log = std::log(o.power);
        k.has_smooth_power_log = std::isfinite(k.smooth_power_log) && k.smooth_power_log != 0.0;
    }
    if (k.has_constant_bailout && k.smooth_bailout_radius > 1.0 && std::isfinite(k.smooth_bailout_radius)) {
        k.smooth_bailout_log = std::log(k.smooth_bailout_radius);
        k.has_smooth_bailout_log = std::isfinite(k.smooth_bailout_log) && k.smooth_bailout_log != 0.0;
    }
    const double norm = selected_bailout_norm(o);
    if (norm > 0.0 && std::isfinite(norm)) {
        k.smooth_norm_exponent = 1.0 / norm;
        k.has_smooth_norm_exponent = 1;
    } else if (std::isinf(norm) && norm > 0.0) {
        k.smooth_norm_exponent = 1.0;
        k.has_smooth_norm_exponent = 1;
    }
    k.texture_active = planes != nullptr && planes->texture_rgb != nullptr && o.texture_enabled != 0;
    const uint32_t offset = static_cast<uint32_t>(offsetof(FractalKfpOptions, show_glitches));
    k.show_glitches = o.struct_size < offset + sizeof(o.show_glitches) ? 1 : (o.show_glitches != 0);
    k.phase_color_strength = o.phase_color_strength;
    k.texture_merge = o.texture_merge;
    k.texture_power = o.texture_power;
    k.texture_ratio = o.texture_ratio;
    if (planes && k.texture_active) {
        k.texture_width = planes->texture_width;
        k.texture_height = planes->texture_height;
        k.texture_stride = planes->texture_stride;
    }
}

}  // namespace kfp
}  // namespace fav
