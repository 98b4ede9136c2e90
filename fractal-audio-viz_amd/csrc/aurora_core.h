// aurora_core.h -- the ordinary (Aurora) palette and the per-frame
// crop / atlas compositors that sit after the field (renderer.cpp:4782-4923,
// 5053-5074, 5146-5300, 13771-13825, 16210-16290, 18399-19566): scalar and
// RGB reprojection of a parent tile and a centred child tile, the seam
// blend, and the palette lookup, one output pixel per lane.
//
// The palette (at most 65536 RGB entries, cos/sin of the host's libm) and
// the bilinear maps are built on the host once per call; the per-pixel
// arithmetic below follows the reference's float / double operation order so
// every byte and every float matches.
#pragma once

#include <cstdint>
#include <vector>

#include "atlas_core.h"

namespace fav {
namespace aurora {

enum Mode : int32_t {
    kColour = 0,       // field -> palette (fractal_colourise[_accents])
    kAccents = 1,      // RGB -> three accents (fractal_apply_aurora_accents)
    kCropRgb = 2,      // fractal_crop_rgb
    kAtlasRgb = 3,     // fractal_atlas_composite_rgb
    kCropField = 4,    // fractal_crop_field
    kAtlasField = 5,   // fractal_atlas_field[_ex]
    kCropColour = 6,   // fractal_crop_colourise[_interior|_accents]
    kAtlasColour = 7,  // fractal_atlas_colourise[_interior|_accents]
};

// Launch parameters (device pointers).
struct Params {
    int32_t mode;
    int32_t width, height;  // output
    const float* parent_field;  // scalar sources (the field itself for kColour)
    const float* child_field;
    const uint8_t* parent_rgb;  // RGB sources (kCropRgb, kAtlasRgb; kAccents reads out_rgb)
    const uint8_t* child_rgb;
    int32_t parent_width, child_width;
    int32_t parent_max_iter, child_max_iter;
    atlas::AxisDev parent_x, parent_y, child_x, child_y;
    const float* edge_x;  // separable seam weights of the child columns / rows
    const float* edge_y;
    const uint8_t* palette;  // palette_size RGB entries
    int32_t palette_size;
    int32_t palette_iter;    // the source_max_iter of the palette lookup
    float index_scale;       // (palette_size - 1) / palette_iter in float
    int32_t has_interior;
    int32_t interior[3];
    int32_t use_child, full_child, child_left, child_top, visible_width, visible_height, feather;
    double parent_bias, child_bias, output_bias;
    int32_t effective_iter;
    double accents[9];
    int32_t rotate;
    double hue_cos, hue_sin;
    float* out_field;
    uint8_t* out_rgb;
};

// ---- host: the palette --------------------------------------------------

// rotate_hue_rgb, renderer.cpp:4818-4839 (YIQ chroma rotation).
FAV_HD void rotate_hue(double& r, double& g, double& b, double c, double s) {
    const double y = 0.299 * r + 0.587 * g + 0.114 * b;
    const double i = 0.596 * r - 0.275 * g - 0.321 * b;
    const double q = 0.212 * r - 0.523 * g + 0.311 * b;
    const double ri = i * c - q * s;
    const double rq = i * s + q * c;
    r = y + 0.956 * ri + 0.621 * rq;
    g = y - 0.272 * ri - 0.647 * rq;
    b = y - 1.106 * ri + 1.703 * rq;
}

// pitch_hue_angle, renderer.cpp:4813-4816.
inline double pitch_hue_angle(double pitch) {
    double signed_pitch = 2.0 * (pitch - 0.5);
    signed_pitch = signed_pitch < -1.0 ? -1.0 : (signed_pitch > 1.0 ? 1.0 : signed_pitch);
    return signed_pitch * 0.16 * 6.283185307179586476925286766559005768;
}

// aurora_palette_for (renderer.cpp:4841-4923) with colour_basis_for
// (4788-4805): palette_size = clamp(4 * max_iter, 4096, 65536) RGB entries.
void build_palette(int max_iter, double phase, double vocal, double pitch, const uint8_t* accents,
                   std::vector<uint8_t>& rgb);

// ---- device / host: per-pixel pieces ------------------------------------

// colour_byte, renderer.cpp:4807-4809 (truncation after the clamp).
FAV_HD uint8_t colour_byte(double v) { return static_cast<uint8_t>(v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v)); }

// rounded_colour_byte, renderer.cpp:5237-5242 (nearest, ties to even).
FAV_HD uint8_t rounded_byte(double v) {
    v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
    int i = static_cast<int>(__builtin_rint(v));
    i = i < 0 ? 0 : (i > 255 ? 255 : i);
    return static_cast<uint8_t>(i);
}

// write_colour_pixel, renderer.cpp:5210-5235.
FAV_HD void colour_pixel(const Params& p, float smooth, uint8_t* d) {
    if (!__builtin_isfinite(smooth) || smooth >= static_cast<float>(p.palette_iter)) {
        d[0] = p.has_interior ? static_cast<uint8_t>(p.interior[0]) : 0;
        d[1] = p.has_interior ? static_cast<uint8_t>(p.interior[1]) : 0;
        d[2] = p.has_interior ? static_cast<uint8_t>(p.interior[2]) : 0;
        return;
    }
    const double scaled = static_cast<double>(smooth) * static_cast<double>(p.index_scale);
    const int index = scaled >= static_cast<double>(p.palette_size - 1)
        ? p.palette_size - 1
        : (!__builtin_isfinite(scaled) || scaled <= 0.0 ? 0 : static_cast<int>(scaled));
    const uint8_t* c = p.palette + static_cast<size_t>(index) * 3;
    d[0] = c[0];
    d[1] = c[1];
    d[2] = c[2];
}

// sample_bilinear_mapped, renderer.cpp:5053-5074 (float arithmetic).
FAV_HD float sample_plain(const float* source, int32_t width, const atlas::AxisDev& ax, const atlas::AxisDev& ay,
                          int x, int y) {
    const int x0 = ax.index0[x], x1 = ax.index1[x];
    const float wx = ax.weight[x];
    const int y0 = ay.index0[y], y1 = ay.index1[y];
    const float wy = ay.weight[y];
    const float* top = source + static_cast<size_t>(y0) * static_cast<size_t>(width);
    const float* bottom = source + static_cast<size_t>(y1) * static_cast<size_t>(width);
    const float t = top[x0] * (1.0F - wx) + top[x1] * wx;
    const float b = bottom[x0] * (1.0F - wx) + bottom[x1] * wx;
    return t * (1.0F - wy) + b * wy;
}

// sample_rgb_bilinear_mapped, renderer.cpp:5262-5300.
FAV_HD void sample_rgb(const uint8_t* source, int32_t width, const atlas::AxisDev& ax, const atlas::AxisDev& ay,
                       int x, int y, uint8_t* d) {
    const int x0 = ax.index0[x], x1 = ax.index1[x];
    const float wx = ax.weight[x];
    const int y0 = ay.index0[y], y1 = ay.index1[y];
    const float wy = ay.weight[y];
    const size_t w = static_cast<size_t>(width);
    const uint8_t* tl = source + (static_cast<size_t>(y0) * w + static_cast<size_t>(x0)) * 3;
    const uint8_t* tr = source + (static_cast<size_t>(y0) * w + static_cast<size_t>(x1)) * 3;
    const uint8_t* bl = source + (static_cast<size_t>(y1) * w + static_cast<size_t>(x0)) * 3;
    const uint8_t* br = source + (static_cast<size_t>(y1) * w + static_cast<size_t>(x1)) * 3;
    const float ix = 1.0F - wx, iy = 1.0F - wy;
    for (int c = 0; c < 3; ++c) {
        const float t = static_cast<float>(tl[c]) * ix + static_cast<float>(tr[c]) * wx;
        const float b = static_cast<float>(bl[c]) * ix + static_cast<float>(br[c]) * wx;
        d[c] = rounded_byte(static_cast<double>(t * iy + b * wy));
    }
}

FAV_HD bool in_child(const Params& p, int x, int y) {
    return p.use_child && x >= p.child_left && x < p.child_left + p.visible_width && y >= p.child_top
        && y < p.child_top + p.visible_height;
}

// fractal_apply_aurora_accents, renderer.cpp:16250-16300 (in place).
FAV_HD void accents_pixel(const Params& p, uint8_t* d) {
    double w[3];
    for (int c = 0; c < 3; ++c) {
        const double v = static_cast<double>(d[c]) / 140.0;
        w[c] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    }
    double r = (w[0] * p.accents[0] + w[1] * p.accents[3] + w[2] * p.accents[6]) / 1.8;
    double g = (w[0] * p.accents[1] + w[1] * p.accents[4] + w[2] * p.accents[7]) / 1.8;
    double b = (w[0] * p.accents[2] + w[1] * p.accents[5] + w[2] * p.accents[8]) / 1.8;
    if (p.rotate) rotate_hue(r, g, b, p.hue_cos, p.hue_sin);
    d[0] = rounded_byte(r);
    d[1] = rounded_byte(g);
    d[2] = rounded_byte(b);
}

// fractal_atlas_composite_rgb's pixel, renderer.cpp:18567-18614.
FAV_HD void atlas_rgb_pixel(const Params& p, int x, int y, uint8_t* d) {
    sample_rgb(p.parent_rgb, p.parent_width, p.parent_x, p.parent_y, x, y, d);
    if (!in_child(p, x, y)) return;
    uint8_t child[3];
    const int cx = x - p.child_left, cy = y - p.child_top;
    sample_rgb(p.child_rgb, p.child_width, p.child_x, p.child_y, cx, cy, child);
    double alpha = 1.0;
    if (p.feather >= 2) {
        int edge = cx < p.visible_width - 1 - cx ? cx : p.visible_width - 1 - cx;
        const int edge_y = cy < p.visible_height - 1 - cy ? cy : p.visible_height - 1 - cy;
        edge = edge < edge_y ? edge : edge_y;
        double linear = static_cast<double>(edge) / static_cast<double>(p.feather);
        linear = linear < 0.0 ? 0.0 : (linear > 1.0 ? 1.0 : linear);
        alpha = linear * linear * (3.0 - 2.0 * linear);
    }
    for (int c = 0; c < 3; ++c)
        d[c] = rounded_byte(static_cast<double>(d[c]) * (1.0 - alpha) + static_cast<double>(child[c]) * alpha);
}

// fractal_atlas_field_ex's pixel, renderer.cpp:18821-18876.
FAV_HD float atlas_field_pixel(const Params& p, int x, int y) {
    const float interior_value = static_cast<float>(static_cast<double>(p.effective_iter) - p.output_bias);
    const float bias = static_cast<float>(p.output_bias);
    bool inside = false;
    if (p.full_child) {
        const float s = atlas::sample(p.child_field, p.child_width, p.child_x, p.child_y, x, y, p.child_max_iter,
                                      inside, p.child_bias);
        return inside ? interior_value : s - bias;
    }
    const float s = atlas::sample(p.parent_field, p.parent_width, p.parent_x, p.parent_y, x, y, p.parent_max_iter,
                                  inside, p.parent_bias);
    float value = inside ? interior_value : s - bias;
    if (in_child(p, x, y)) {
        const int cx = x - p.child_left, cy = y - p.child_top;
        bool child_inside = false;
        const float cs = atlas::sample(p.child_field, p.child_width, p.child_x, p.child_y, cx, cy, p.child_max_iter,
                                       child_inside, p.child_bias);
        const float child_value = child_inside ? interior_value : cs - bias;
        float alpha = 1.0F;
        if (p.feather >= 2) alpha = p.edge_x[cx] < p.edge_y[cy] ? p.edge_x[cx] : p.edge_y[cy];
        value = value * (1.0F - alpha) + child_value * alpha;
    }
    return value;
}

// crop_colourise_impl's pixel, renderer.cpp:18964-19000.
FAV_HD void crop_colour_pixel(const Params& p, int x, int y, uint8_t* d) {
    bool inside = false;
    const float s = atlas::sample(p.parent_field, p.parent_width, p.parent_x, p.parent_y, x, y, p.parent_max_iter,
                                  inside, 0.0);
    if (inside) {
        d[0] = p.has_interior ? static_cast<uint8_t>(p.interior[0]) : 0;
        d[1] = p.has_interior ? static_cast<uint8_t>(p.interior[1]) : 0;
        d[2] = p.has_interior ? static_cast<uint8_t>(p.interior[2]) : 0;
        return;
    }
    const double scaled = static_cast<double>(s) * static_cast<double>(p.index_scale);
    const int index = scaled >= static_cast<double>(p.palette_size - 1)
        ? p.palette_size - 1
        : (!__builtin_isfinite(scaled) || scaled <= 0.0 ? 0 : static_cast<int>(scaled));
    const uint8_t* c = p.palette + static_cast<size_t>(index) * 3;
    d[0] = c[0];
    d[1] = c[1];
    d[2] = c[2];
}

// atlas_colourise_impl's pixel, renderer.cpp:19265-19395.
FAV_HD void atlas_colour_pixel(const Params& p, int x, int y, uint8_t* d) {
    const float interior_value = static_cast<float>(p.palette_iter);
    bool inside = false;
    if (p.full_child) {
        const float s = atlas::sample(p.child_field, p.child_width, p.child_x, p.child_y, x, y, p.child_max_iter,
                                      inside, 0.0);
        colour_pixel(p, inside ? interior_value : s, d);
        return;
    }
    if (!in_child(p, x, y)) {
        const float s = atlas::sample(p.parent_field, p.parent_width, p.parent_x, p.parent_y, x, y,
                                      p.parent_max_iter, inside, 0.0);
        colour_pixel(p, inside ? interior_value : s, d);
        return;
    }
    const int cx = x - p.child_left, cy = y - p.child_top;
    bool child_inside = false;
    const float cs = atlas::sample(p.child_field, p.child_width, p.child_x, p.child_y, cx, cy, p.child_max_iter,
                                   child_inside, 0.0);
    const float alpha = p.feather >= 2 ? (p.edge_x[cx] < p.edge_y[cy] ? p.edge_x[cx] : p.edge_y[cy]) : 1.0F;
    if (alpha >= 0.999999F) {
        colour_pixel(p, child_inside ? interior_value : cs, d);
        return;
    }
    const float ps = atlas::sample(p.parent_field, p.parent_width, p.parent_x, p.parent_y, x, y, p.parent_max_iter,
                                   inside, 0.0);
    uint8_t parent_rgb[3], child_rgb[3];
    colour_pixel(p, inside ? interior_value : ps, parent_rgb);
    colour_pixel(p, child_inside ? interior_value : cs, child_rgb);
    const double a = static_cast<double>(alpha);
    for (int c = 0; c < 3; ++c)
        d[c] = rounded_byte(static_cast<double>(parent_rgb[c]) * (1.0 - a) + static_cast<double>(child_rgb[c]) * a);
}

}  // namespace aurora
}  // namespace fav
