// atlas_core.h -- the raw KFP atlas compose step (fractal_atlas_colourise_kfp_raw,
// renderer.cpp:17969-18265): the parent tile and the centred child tile
// reprojected into one frame-sized scalar field, which the KFP colour kernel
// then colours in one pass, so its screen-space difference and slope
// stencils stay continuous across the hand-over seam.
//
// The bilinear maps (one entry per output column / row) are built on the
// host with the reference's double arithmetic; the per-pixel sampling and the
// seam blend run on the device, one pixel per lane.
#pragma once

#include <cstdint>
#include <vector>

#include "fav_numeric.h"

namespace fav {
namespace atlas {

// A bilinear map of one axis: source indices and the float weight of the
// second (BilinearAxis, renderer.cpp:4925-4929).
struct Axis {
    std::vector<int32_t> index0, index1;
    std::vector<float> weight;
};

// fill_bilinear_axis, renderer.cpp:4931-4962: the centred crop of a tile
// zoomed by zoom_factor, resampled to destination_size pixels.
inline void fill_axis(Axis& axis, int source_size, int destination_size, double zoom_factor) {
    zoom_factor = zoom_factor > 1.0 ? zoom_factor : 1.0;
    const double inverse_zoom = 1.0 / zoom_factor;
    const double crop_size = static_cast<double>(source_size) * inverse_zoom;
    const double left = (static_cast<double>(source_size) - crop_size) * 0.5;
    const double step = crop_size / static_cast<double>(destination_size);
    axis.index0.resize(static_cast<size_t>(destination_size));
    axis.index1.resize(static_cast<size_t>(destination_size));
    axis.weight.resize(static_cast<size_t>(destination_size));
    for (int d = 0; d < destination_size; ++d) {
        double source = left + (static_cast<double>(d) + 0.5) * step - 0.5;
        source = source < 0.0 ? 0.0 : (source > static_cast<double>(source_size - 1) ? source_size - 1.0 : source);
        const int i0 = static_cast<int>(std::floor(source));
        axis.index0[static_cast<size_t>(d)] = i0;
        axis.index1[static_cast<size_t>(d)] = i0 + 1 < source_size - 1 ? i0 + 1 : source_size - 1;
        axis.weight[static_cast<size_t>(d)] = static_cast<float>(source - static_cast<double>(i0));
    }
}

// fill_bilinear_axis_window, renderer.cpp:4969-5002: the same crop mapped
// onto a window that starts `origin` pixels before the logical destination
// of `span` pixels (the child's stencil halo).
inline void fill_axis_window(Axis& axis, int source_size, int destination_size, double origin, double span,
                             double zoom) {
    zoom = zoom > 1.0 ? zoom : 1.0;
    const double crop_size = static_cast<double>(source_size) / zoom;
    const double source_left = (static_cast<double>(source_size) - crop_size) * 0.5;
    axis.index0.resize(static_cast<size_t>(destination_size));
    axis.index1.resize(static_cast<size_t>(destination_size));
    axis.weight.resize(static_cast<size_t>(destination_size));
    for (int d = 0; d < destination_size; ++d) {
        const double logical = origin + static_cast<double>(d) + 0.5;
        double source = source_left + logical * crop_size / span - 0.5;
        source = source < 0.0 ? 0.0 : (source > static_cast<double>(source_size - 1) ? source_size - 1.0 : source);
        const int i0 = static_cast<int>(std::floor(source));
        axis.index0[static_cast<size_t>(d)] = i0;
        axis.index1[static_cast<size_t>(d)] = i0 + 1 < source_size - 1 ? i0 + 1 : source_size - 1;
        axis.weight[static_cast<size_t>(d)] = static_cast<float>(source - static_cast<double>(i0));
    }
}

// kfp_atlas_seam_feather, renderer.cpp:99-110 (at most 64 pixels, half the
// smaller child side, and never the whole child).
inline int seam_feather(int width, int height) {
    const int minimum = width < height ? width : height;
    if (minimum <= 0) return 0;
    int requested = minimum / 2 < 64 ? minimum / 2 : 64;
    if (requested < 2) requested = 2;
    const int plateau = (minimum - 1) / 2 > 0 ? (minimum - 1) / 2 : 0;
    return requested < plateau ? requested : plateau;
}

// Device view of one axis.
struct AxisDev {
    const int32_t* index0;
    const int32_t* index1;
    const float* weight;
};

// Launch parameters of the compose kernel (device pointers).
struct ComposeParams {
    const float* parent;
    const float* child;
    AxisDev parent_x, parent_y, child_x, child_y;
    float* out;
    double parent_bias, child_bias, field_bias;
    int32_t parent_width, parent_max_iter, child_width, child_max_iter;
    int32_t width, height, max_iter;
    int32_t use_child, visible_left, visible_top, visible_width, visible_height;
    int32_t child_left, child_top, feather;
};

// sample_bilinear_mapped_preserving_interior, renderer.cpp:5146-5208: a
// bilinear average of the exterior corners; an interior (cap) weight of at
// least one half makes the sample interior.
FAV_HD float sample(const float* source, int32_t width, const AxisDev& ax, const AxisDev& ay, int x, int y,
                    int32_t max_iter, bool& inside, double bias) {
    const int x0 = ax.index0[x], x1 = ax.index1[x];
    const double wx = static_cast<double>(ax.weight[x]);
    const int y0 = ay.index0[y], y1 = ay.index1[y];
    const double wy = static_cast<double>(ay.weight[y]);
    const float* top = source + static_cast<size_t>(y0) * static_cast<size_t>(width);
    const float* bottom = source + static_cast<size_t>(y1) * static_cast<size_t>(width);
    const float values[4] = {top[x0], top[x1], bottom[x0], bottom[x1]};
    const double weights[4] = {(1.0 - wy) * (1.0 - wx), (1.0 - wy) * wx, wy * (1.0 - wx), wy * wx};
    double interior = 0.0, exterior = 0.0, value = 0.0;
    for (int k = 0; k < 4; ++k) {
        const float v = values[k];
        if (__builtin_isfinite(v) && static_cast<double>(v) >= static_cast<double>(max_iter) - bias) {
            interior += weights[k];
        } else if (__builtin_isfinite(v)) {
            exterior += weights[k];
            value += (static_cast<double>(v) + bias) * weights[k];
        }
    }
    if (interior >= 0.5) {
        inside = true;
        return static_cast<float>(max_iter);
    }
    inside = false;
    if (exterior <= 1.0e-12) return 0.0F;  // no finite neighbour
    return static_cast<float>(value / exterior);
}

// encode_render_value, renderer.cpp:119-125, in double: exact for the
// integer biases of the atlas encoding (a float minus an integer below 2^31
// needs at most 52 significant bits).
FAV_HD float encode(double value, double bias) { return static_cast<float>(value - bias); }

// One output pixel of the composed field (renderer.cpp:18085-18250).
FAV_HD float compose_pixel(const ComposeParams& a, int x, int y) {
    bool inside = false;
    const float smooth = sample(a.parent, a.parent_width, a.parent_x, a.parent_y, x, y, a.parent_max_iter, inside,
                                a.parent_bias);
    float value = inside ? encode(static_cast<double>(a.max_iter), a.field_bias)
                         : encode(static_cast<double>(smooth), a.field_bias);
    if (a.use_child && x >= a.visible_left && x < a.visible_left + a.visible_width && y >= a.visible_top
        && y < a.visible_top + a.visible_height) {
        bool child_inside = false;
        const float child_smooth = sample(a.child, a.child_width, a.child_x, a.child_y, x - a.child_left,
                                          y - a.child_top, a.child_max_iter, child_inside, a.child_bias);
        const float child_value = child_inside ? encode(static_cast<double>(a.max_iter), a.field_bias)
                                               : encode(static_cast<double>(child_smooth), a.field_bias);
        float alpha = 1.0F;
        if (a.feather >= 2) {  // smoothstep over the feather width from the child's edge
            const int vx = x - a.visible_left, vy = y - a.visible_top;
            int edge = vx < a.visible_width - 1 - vx ? vx : a.visible_width - 1 - vx;
            const int edge_y = vy < a.visible_height - 1 - vy ? vy : a.visible_height - 1 - vy;
            if (edge_y < edge) edge = edge_y;
            float linear = static_cast<float>(edge) / static_cast<float>(a.feather);
            linear = linear < 0.0F ? 0.0F : (linear > 1.0F ? 1.0F : linear);
            alpha = linear * linear * (3.0F - 2.0F * linear);
        }
        value = alpha >= 0.999999F ? child_value : value * (1.0F - alpha) + child_value * alpha;
    }
    return value;
}

// ---- the plane-aware compose (fractal_atlas_colourise_kfp_planes,
// renderer.cpp:17323-17793): the scalar field and every Kalles plane that
// either tile carries, reprojected into the frame in one pass per pixel.

// One source tile of the plane-aware compose (device pointers; absent
// planes are null).
struct PlaneSource {
    const float* field;
    int32_t width, max_iter;
    double bias;  // centred storage: the tile's own cap
    AxisDev x, y;
    const int64_t* ints[2];      // orbit_iteration, iteration
    const double* doubles[7];    // bailout, transition, phase, de_x, de_y, test1, test2
};

struct PlanesCompose {
    PlaneSource primary, secondary;
    int32_t use_secondary, visible_left, visible_top, visible_width, visible_height, feather;
    int32_t width, height, target_max_iter;
    double field_bias;  // the KFP options' bias of the composed field
    float* out_field;
    int64_t* out_ints[2];
    double* out_doubles[7];
};

// (preserve_negative, reset_inside) of each double plane (renderer.cpp:17626-17651).
FAV_HD bool plane_preserves_negative(int k) { return k == 1; }
FAV_HD bool plane_resets_inside(int k) { return k != 0; }

// sample_nearest_mapped_*, renderer.cpp:5105-5139.
template <typename T>
FAV_HD T sample_nearest(const T* source, int32_t width, const AxisDev& ax, const AxisDev& ay, int x, int y) {
    const int xi = ax.weight[x] < 0.5F ? ax.index0[x] : ax.index1[x];
    const int yi = ay.weight[y] < 0.5F ? ay.index0[y] : ay.index1[y];
    return source[static_cast<size_t>(yi) * static_cast<size_t>(width) + static_cast<size_t>(xi)];
}

// sample_bilinear_mapped_double, renderer.cpp:5076-5103, then the
// negative-nearest override of a sign-carrying plane (17588-17597).
FAV_HD double sample_plane(const double* source, int32_t width, const AxisDev& ax, const AxisDev& ay, int x, int y,
                           bool preserve_negative) {
    const int x0 = ax.index0[x], x1 = ax.index1[x];
    const double wx = static_cast<double>(ax.weight[x]);
    const int y0 = ay.index0[y], y1 = ay.index1[y];
    const double wy = static_cast<double>(ay.weight[y]);
    const double* top = source + static_cast<size_t>(y0) * static_cast<size_t>(width);
    const double* bottom = source + static_cast<size_t>(y1) * static_cast<size_t>(width);
    auto ordinary = [preserve_negative](double v) { return preserve_negative && v < 0.0 ? 0.0 : v; };
    const double t = ordinary(top[x0]) * (1.0 - wx) + ordinary(top[x1]) * wx;
    const double b = ordinary(bottom[x0]) * (1.0 - wx) + ordinary(bottom[x1]) * wx;
    double value = t * (1.0 - wy) + b * wy;
    if (preserve_negative) {
        const double nearest = sample_nearest(source, width, ax, ay, x, y);
        if (nearest < 0.0) value = nearest;
    }
    return value;
}

FAV_HD void compose_planes_pixel(const PlanesCompose& a, int x, int y) {
    const size_t i = static_cast<size_t>(y) * static_cast<size_t>(a.width) + static_cast<size_t>(x);
    const float interior = encode(static_cast<double>(a.target_max_iter), a.field_bias);
    const PlaneSource& p = a.primary;
    bool inside = false;
    const float smooth = sample(p.field, p.width, p.x, p.y, x, y, p.max_iter, inside, p.bias);
    float value = inside ? interior : encode(static_cast<double>(smooth), a.field_bias);
    int64_t ints[2];
    double doubles[7];
    for (int k = 0; k < 2; ++k) {
        ints[k] = 0;
        if (p.ints[k])
            ints[k] = inside ? static_cast<int64_t>(a.target_max_iter) : sample_nearest(p.ints[k], p.width, p.x, p.y, x, y);
    }
    for (int k = 0; k < 7; ++k) {
        doubles[k] = 0.0;
        if (p.doubles[k]) {
            const double v = sample_plane(p.doubles[k], p.width, p.x, p.y, x, y, plane_preserves_negative(k));
            doubles[k] = inside && plane_resets_inside(k) ? 0.0 : v;
        }
    }
    if (a.use_secondary && x >= a.visible_left && x < a.visible_left + a.visible_width && y >= a.visible_top
        && y < a.visible_top + a.visible_height) {
        const PlaneSource& c = a.secondary;
        const int cx = x - a.visible_left, cy = y - a.visible_top;
        bool child_inside = false;
        const float cs = sample(c.field, c.width, c.x, c.y, cx, cy, c.max_iter, child_inside, c.bias);
        const float child_value = child_inside ? interior : encode(static_cast<double>(cs), a.field_bias);
        float alpha = 1.0F;
        if (a.feather >= 2) {  // renderer.cpp:17675-17689
            int edge = cx < a.visible_width - 1 - cx ? cx : a.visible_width - 1 - cx;
            const int edge_y = cy < a.visible_height - 1 - cy ? cy : a.visible_height - 1 - cy;
            edge = edge < edge_y ? edge : edge_y;
            float linear = static_cast<float>(edge) / static_cast<float>(a.feather);
            linear = linear < 0.0F ? 0.0F : (linear > 1.0F ? 1.0F : linear);
            alpha = linear * linear * (3.0F - 2.0F * linear);
        }
        value = value * (1.0F - alpha) + child_value * alpha;
        for (int k = 0; k < 2; ++k)
            if (c.ints[k] && a.out_ints[k])
                ints[k] = child_inside ? static_cast<int64_t>(a.target_max_iter)
                                       : sample_nearest(c.ints[k], c.width, c.x, c.y, cx, cy);
        for (int k = 0; k < 7; ++k) {
            if (!c.doubles[k] || !a.out_doubles[k]) continue;
            const bool negative = plane_preserves_negative(k);
            double v = sample_plane(c.doubles[k], c.width, c.x, c.y, cx, cy, negative);
            if (child_inside && plane_resets_inside(k)) v = 0.0;
            if (!p.doubles[k]) {
                doubles[k] = v;
                continue;
            }
            const double parent_value = doubles[k];
            const double blended = parent_value * (1.0 - static_cast<double>(alpha)) + v * static_cast<double>(alpha);
            if (negative) {
                const double selected = alpha >= 0.5F ? v : parent_value;
                doubles[k] = selected < 0.0 ? selected : blended;
            } else {
                doubles[k] = blended;
            }
        }
    }
    a.out_field[i] = value;
    for (int k = 0; k < 2; ++k)
        if (a.out_ints[k]) a.out_ints[k][i] = ints[k];
    for (int k = 0; k < 7; ++k)
        if (a.out_doubles[k]) a.out_doubles[k][i] = doubles[k];
}

}  // namespace atlas
}  // namespace fav
