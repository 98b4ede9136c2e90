// aurora_abi.cpp -- C ABI of the ordinary-palette colouriser and the
// per-frame crop / atlas compositors (renderer.h:400-449, 713-810,
// 943-1067): the reference's argument checks and messages, the palette, the
// bilinear maps and the seam weights on the host, the per-pixel pass on the
// device (aurora_core.h, aurora_kernel.hip).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fractal_hip.h"
#include "atlas_core.h"
#include "aurora_core.h"
#include "fav_render.h"

namespace {

using fav::AuroraInputs;
using fav::atlas::Axis;
namespace aurora = fav::aurora;

constexpr int kMaxIterations = 10'000'000;  // renderer.cpp:88
constexpr uint64_t kMaxPixels = 100'000'000;  // renderer.cpp:91
constexpr int kMaxThreads = 4096;

bool valid_dimensions(int w, int h) {  // renderer.cpp:170-175
    return w > 0 && h > 0 && static_cast<uint64_t>(w) <= kMaxPixels / static_cast<uint64_t>(h);
}
bool valid_iterations(int n) { return n > 0 && n <= kMaxIterations; }
bool valid_threads(int n) { return n >= 0 && n <= kMaxThreads; }
bool valid_controls(double phase, double vocal, double instrumental, double pitch) {  // renderer.cpp:177-187
    return std::isfinite(phase) && std::isfinite(vocal) && std::isfinite(instrumental) && std::isfinite(pitch);
}
bool valid_interior(const int* c) {
    return !c || (c[0] >= 0 && c[0] <= 255 && c[1] >= 0 && c[1] <= 255 && c[2] >= 0 && c[2] <= 255);
}

template <typename Fn>
int guarded(const char* unknown, Fn body) {
    try {
        body();
        fav::set_last_error("");
        return 0;
    } catch (const std::exception& e) {
        fav::set_last_error(e.what());
        return 1;
    } catch (...) {
        fav::set_last_error(unknown);
        return 1;
    }
}

aurora::Params base_params(int mode, int width, int height) {
    aurora::Params p{};
    p.mode = mode;
    p.width = width;
    p.height = height;
    return p;
}

void set_palette(aurora::Params& p, const std::vector<uint8_t>& palette, int palette_iter, const int* interior) {
    p.palette_size = static_cast<int32_t>(palette.size() / 3);
    p.palette_iter = palette_iter;
    p.index_scale = static_cast<float>(p.palette_size - 1) / static_cast<float>(palette_iter);
    p.has_interior = interior != nullptr;
    for (int c = 0; c < 3; ++c) p.interior[c] = interior ? interior[c] : 0;
}

// The visible child rectangle, centred (renderer.cpp:18510-18523).
void child_rectangle(aurora::Params& p, bool use_child, double child_fraction) {
    p.use_child = use_child;
    p.visible_width = use_child
        ? std::max(1, static_cast<int>(std::lround(static_cast<double>(p.width) * child_fraction))) : 0;
    p.visible_height = use_child
        ? std::max(1, static_cast<int>(std::lround(static_cast<double>(p.height) * child_fraction))) : 0;
    p.child_left = (p.width - p.visible_width) / 2;
    p.child_top = (p.height - p.visible_height) / 2;
}

// Separable smoothstep seam weights of the child columns and rows, in float
// (renderer.cpp:18791-18810, 19226-19245).
void seam_weights(int visible_width, int visible_height, int feather, std::vector<float>& ex, std::vector<float>& ey) {
    auto fill = [feather](std::vector<float>& v, int n) {
        v.resize(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            const int edge = std::min(i, n - 1 - i);
            const float linear = std::min(1.0F, static_cast<float>(edge) / static_cast<float>(feather));
            v[static_cast<size_t>(i)] = linear * linear * (3.0F - 2.0F * linear);
        }
    };
    fill(ex, visible_width);
    fill(ey, visible_height);
}

// colourise_field_impl, renderer.cpp:13771-13825.
int colourise_field(const float* field, uint8_t* output, int width, int height, int max_iter, double phase,
                    double vocal, double instrumental, double pitch, int threads, const int* interior,
                    const uint8_t* accents, const char* message) {
    return guarded("native colouriser failed with an unknown exception", [&] {
        if (!field || !output || !valid_dimensions(width, height) || !valid_iterations(max_iter)
            || !valid_threads(threads) || !valid_controls(phase, vocal, instrumental, pitch)
            || !valid_interior(interior))
            throw std::runtime_error(message);
        std::vector<uint8_t> palette;
        aurora::build_palette(max_iter, phase, vocal, pitch, accents, palette);
        aurora::Params p = base_params(aurora::kColour, width, height);
        set_palette(p, palette, max_iter, interior);
        AuroraInputs in;
        in.parent_field = field;
        in.parent_field_count = static_cast<size_t>(width) * static_cast<size_t>(height);
        in.palette = &palette;
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

// crop_colourise_impl, renderer.cpp:18930-19020.
int crop_colourise(const float* source, int source_width, int source_height, uint8_t* output, int output_width,
                   int output_height, double zoom_factor, int max_iter, double phase, double vocal,
                   double instrumental, double pitch, int threads, const int* interior, const uint8_t* accents) {
    return guarded("native crop colouriser failed with an unknown exception", [&] {
        if (!source || !output || !valid_dimensions(source_width, source_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(max_iter)
            || !valid_threads(threads) || !std::isfinite(zoom_factor) || zoom_factor <= 0.0
            || !valid_controls(phase, vocal, instrumental, pitch) || !valid_interior(interior))
            throw std::runtime_error("invalid native crop/colour dimensions or palette");
        zoom_factor = std::max(zoom_factor, 1.0);
        std::vector<uint8_t> palette;
        aurora::build_palette(max_iter, phase, vocal, pitch, accents, palette);
        aurora::Params p = base_params(aurora::kCropColour, output_width, output_height);
        set_palette(p, palette, max_iter, interior);
        p.parent_width = source_width;
        p.parent_max_iter = max_iter;
        Axis ax, ay;
        fav::atlas::fill_axis(ax, source_width, output_width, zoom_factor);
        fav::atlas::fill_axis(ay, source_height, output_height, zoom_factor);
        AuroraInputs in;
        in.parent_field = source;
        in.parent_field_count = static_cast<size_t>(source_width) * static_cast<size_t>(source_height);
        in.axes[0] = &ax;
        in.axes[1] = &ay;
        in.palette = &palette;
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

// atlas_colourise_impl, renderer.cpp:19137-19410.
int atlas_colourise(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                    const float* child, int child_width, int child_height, int child_max_iter, uint8_t* output,
                    int output_width, int output_height, double parent_zoom, double child_fraction,
                    int palette_max_iter, double phase, double vocal, double instrumental, double pitch,
                    int threads, const int* interior, const uint8_t* accents) {
    return guarded("native atlas colouriser failed with an unknown exception", [&] {
        if (!parent || !output || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(parent_max_iter)
            || !valid_threads(threads) || !std::isfinite(parent_zoom) || parent_zoom <= 0.0
            || !std::isfinite(child_fraction) || child_fraction < 0.0 || child_fraction > 1.0
            || palette_max_iter < 0 || palette_max_iter > kMaxIterations
            || !valid_controls(phase, vocal, instrumental, pitch) || !valid_interior(interior))
            throw std::runtime_error("invalid native atlas dimensions or controls");
        const bool use_child = child != nullptr && child_fraction > 0.0;
        if (use_child && (!valid_dimensions(child_width, child_height) || !valid_iterations(child_max_iter)))
            throw std::runtime_error("invalid native atlas child tile");
        const int effective_child_iter = use_child ? child_max_iter : 0;
        const int effective = std::max(1, std::max(palette_max_iter, std::max(parent_max_iter, effective_child_iter)));
        std::vector<uint8_t> palette;
        aurora::build_palette(effective, phase, vocal, pitch, accents, palette);
        aurora::Params p = base_params(aurora::kAtlasColour, output_width, output_height);
        set_palette(p, palette, effective, interior);
        child_rectangle(p, use_child, child_fraction);
        p.feather = use_child ? std::min(48, std::min(p.visible_width / 8, p.visible_height / 8)) : 0;
        p.full_child = use_child && child_fraction >= 0.999999;
        p.parent_width = parent_width;
        p.parent_max_iter = parent_max_iter;
        p.child_width = use_child ? child_width : 0;
        p.child_max_iter = use_child ? child_max_iter : 0;
        Axis axes[4];
        std::vector<float> ex, ey;
        AuroraInputs in;
        if (!p.full_child) {
            fav::atlas::fill_axis(axes[0], parent_width, output_width, parent_zoom);
            fav::atlas::fill_axis(axes[1], parent_height, output_height, parent_zoom);
            in.parent_field = parent;
            in.parent_field_count = static_cast<size_t>(parent_width) * static_cast<size_t>(parent_height);
            in.axes[0] = &axes[0];
            in.axes[1] = &axes[1];
        }
        if (use_child) {
            fav::atlas::fill_axis(axes[2], child_width, p.full_child ? output_width : p.visible_width, 1.0);
            fav::atlas::fill_axis(axes[3], child_height, p.full_child ? output_height : p.visible_height, 1.0);
            in.child_field = child;
            in.child_field_count = static_cast<size_t>(child_width) * static_cast<size_t>(child_height);
            in.axes[2] = &axes[2];
            in.axes[3] = &axes[3];
        }
        if (use_child && !p.full_child && p.feather >= 2) {
            seam_weights(p.visible_width, p.visible_height, p.feather, ex, ey);
            in.edge_x = &ex;
            in.edge_y = &ey;
        }
        in.palette = &palette;
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

// fractal_atlas_composite_rgb, renderer.cpp:18467-18627.
int atlas_composite_rgb(const uint8_t* parent, int parent_width, int parent_height, const uint8_t* child,
                        int child_width, int child_height, uint8_t* output, int output_width, int output_height,
                        double parent_zoom, double child_fraction, double child_zoom, int feather, int threads) {
    return guarded("native RGB atlas compositor failed with an unknown exception", [&] {
        if (!parent || !output || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || !valid_threads(threads)
            || !std::isfinite(parent_zoom) || parent_zoom <= 0.0 || !std::isfinite(child_fraction)
            || child_fraction < 0.0 || child_fraction > 1.0 || !std::isfinite(child_zoom) || child_zoom <= 0.0
            || feather < 0)
            throw std::runtime_error("invalid native RGB atlas dimensions or controls");
        parent_zoom = std::max(parent_zoom, 1.0);
        child_zoom = std::max(child_zoom, 1.0);
        const bool use_child = child != nullptr && child_fraction > 0.0;
        if (!use_child && child_fraction > 0.0)
            throw std::runtime_error("RGB atlas child data is missing for a positive fraction");
        if (!use_child && child != nullptr) throw std::runtime_error("RGB atlas child needs a positive visible fraction");
        if (!use_child && (child_width != 0 || child_height != 0))
            throw std::runtime_error("invalid native RGB atlas child absence");
        if (use_child
            && (!valid_dimensions(child_width, child_height) || feather > std::min(output_width, output_height)))
            throw std::runtime_error("invalid native RGB atlas child tile");
        aurora::Params p = base_params(aurora::kAtlasRgb, output_width, output_height);
        child_rectangle(p, use_child, child_fraction);
        if (!use_child) p.child_left = p.child_top = 0;
        p.feather = use_child ? std::min(feather, std::min(p.visible_width, p.visible_height)) : 0;
        p.parent_width = parent_width;
        p.child_width = use_child ? child_width : 0;
        Axis axes[4];
        AuroraInputs in;
        fav::atlas::fill_axis(axes[0], parent_width, output_width, parent_zoom);
        fav::atlas::fill_axis(axes[1], parent_height, output_height, parent_zoom);
        in.parent_rgb = parent;
        in.parent_rgb_bytes = static_cast<size_t>(parent_width) * static_cast<size_t>(parent_height) * 3;
        in.axes[0] = &axes[0];
        in.axes[1] = &axes[1];
        if (use_child) {
            fav::atlas::fill_axis(axes[2], child_width, p.visible_width, child_zoom);
            fav::atlas::fill_axis(axes[3], child_height, p.visible_height, child_zoom);
            in.child_rgb = child;
            in.child_rgb_bytes = static_cast<size_t>(child_width) * static_cast<size_t>(child_height) * 3;
            in.axes[2] = &axes[2];
            in.axes[3] = &axes[3];
        }
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

// fractal_atlas_field_ex, renderer.cpp:18697-18888.
int atlas_field_ex(const float* parent, int parent_width, int parent_height, int parent_max_iter, const float* child,
                   int child_width, int child_height, int child_max_iter, float* output, int output_width,
                   int output_height, double parent_zoom, double child_fraction, double child_zoom,
                   double parent_field_bias, double child_field_bias, double output_field_bias, int palette_max_iter,
                   int feather, int threads) {
    return guarded("native scalar atlas compositor failed with an unknown exception", [&] {
        if (!parent || !output || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(parent_max_iter)
            || !valid_threads(threads) || !std::isfinite(parent_zoom) || parent_zoom <= 0.0
            || !std::isfinite(child_fraction) || child_fraction < 0.0 || child_fraction > 1.0
            || !std::isfinite(child_zoom) || child_zoom <= 0.0 || !std::isfinite(parent_field_bias)
            || parent_field_bias < 0.0 || parent_field_bias > static_cast<double>(parent_max_iter)
            || !std::isfinite(child_field_bias) || child_field_bias < 0.0 || !std::isfinite(output_field_bias)
            || output_field_bias < 0.0 || feather < 0 || !valid_iterations(palette_max_iter))
            throw std::runtime_error("invalid native scalar atlas dimensions");
        const bool use_child = child != nullptr && child_fraction > 0.0;
        if (use_child
            && (!valid_dimensions(child_width, child_height) || !valid_iterations(child_max_iter)
                || child_field_bias > static_cast<double>(child_max_iter)))
            throw std::runtime_error("invalid native scalar atlas child tile");
        const int effective = std::max(palette_max_iter, std::max(parent_max_iter, use_child ? child_max_iter : 0));
        if (!valid_iterations(effective) || output_field_bias > static_cast<double>(effective))
            throw std::runtime_error("invalid native scalar atlas iteration cap");
        aurora::Params p = base_params(aurora::kAtlasField, output_width, output_height);
        child_rectangle(p, use_child, child_fraction);
        if (!use_child) p.child_left = p.child_top = 0;
        p.full_child = use_child && child_fraction >= 0.999999;
        p.feather = use_child && !p.full_child
            ? std::min(feather, std::min(p.visible_width / 2, p.visible_height / 2)) : 0;
        p.parent_width = parent_width;
        p.parent_max_iter = parent_max_iter;
        p.child_width = use_child ? child_width : 0;
        p.child_max_iter = use_child ? child_max_iter : 0;
        p.parent_bias = parent_field_bias;
        p.child_bias = child_field_bias;
        p.output_bias = output_field_bias;
        p.effective_iter = effective;
        Axis axes[4];
        std::vector<float> ex, ey;
        AuroraInputs in;
        if (!p.full_child) {
            fav::atlas::fill_axis(axes[0], parent_width, output_width, parent_zoom);
            fav::atlas::fill_axis(axes[1], parent_height, output_height, parent_zoom);
            in.parent_field = parent;
            in.parent_field_count = static_cast<size_t>(parent_width) * static_cast<size_t>(parent_height);
            in.axes[0] = &axes[0];
            in.axes[1] = &axes[1];
        }
        if (use_child) {
            const double zoom = std::max(child_zoom, 1.0);
            fav::atlas::fill_axis(axes[2], child_width, p.full_child ? output_width : p.visible_width, zoom);
            fav::atlas::fill_axis(axes[3], child_height, p.full_child ? output_height : p.visible_height, zoom);
            in.child_field = child;
            in.child_field_count = static_cast<size_t>(child_width) * static_cast<size_t>(child_height);
            in.axes[2] = &axes[2];
            in.axes[3] = &axes[3];
        }
        if (use_child && !p.full_child && p.feather >= 2) {
            seam_weights(p.visible_width, p.visible_height, p.feather, ex, ey);
            in.edge_x = &ex;
            in.edge_y = &ey;
        }
        in.out_field = output;
        fav::aurora_run(p, in);
    });
}

}  // namespace

extern "C" {

int fractal_colourise(const float* field, uint8_t* output, int width, int height, int max_iter, double phase,
                      double vocal, double instrumental, double pitch, int threads) {
    return colourise_field(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, threads,
                           nullptr, nullptr, "invalid native colour dimensions or palette");
}

int fractal_colourise_accents(const float* field, uint8_t* output, int width, int height, int max_iter,
                              double phase, double vocal, double instrumental, double pitch, const uint8_t* accents,
                              int interior_red, int interior_green, int interior_blue, int threads) {
    if (!accents) {
        fav::set_last_error("ordinary palette accents are required");
        return 1;
    }
    const int interior[3] = {interior_red, interior_green, interior_blue};
    return colourise_field(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, threads,
                           interior, accents, "invalid native colour dimensions or palette");
}

int fractal_apply_aurora_accents(uint8_t* output, int width, int height, const uint8_t* accents, double pitch,
                                 int threads) {
    return guarded("Aurora accent pass failed with an unknown exception", [&] {
        if (!output || !accents || !valid_dimensions(width, height) || !valid_threads(threads)
            || !std::isfinite(pitch))
            throw std::runtime_error("invalid Aurora accent dimensions or controls");
        aurora::Params p = base_params(aurora::kAccents, width, height);
        for (int i = 0; i < 9; ++i) p.accents[i] = static_cast<double>(accents[i]);
        const double angle = aurora::pitch_hue_angle(pitch);
        p.rotate = std::abs(angle) > 1.0e-15;
        p.hue_cos = p.rotate ? std::cos(angle) : 1.0;
        p.hue_sin = p.rotate ? std::sin(angle) : 0.0;
        AuroraInputs in;
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

int fractal_crop_rgb(const uint8_t* source, int source_width, int source_height, uint8_t* output, int output_width,
                     int output_height, double zoom_factor, int threads) {
    return guarded("native RGB crop failed with an unknown exception", [&] {
        if (!source || !output || !valid_dimensions(source_width, source_height)
            || !valid_dimensions(output_width, output_height) || !valid_threads(threads)
            || !std::isfinite(zoom_factor) || zoom_factor <= 0.0)
            throw std::runtime_error("invalid native RGB crop dimensions");
        zoom_factor = std::max(zoom_factor, 1.0);
        // The reference copies an identity crop (renderer.cpp:18416-18425);
        // the sampler reproduces it exactly there (index d, weight 0), so
        // the identity runs through the same kernel.
        aurora::Params p = base_params(aurora::kCropRgb, output_width, output_height);
        p.parent_width = source_width;
        Axis ax, ay;
        fav::atlas::fill_axis(ax, source_width, output_width, zoom_factor);
        fav::atlas::fill_axis(ay, source_height, output_height, zoom_factor);
        AuroraInputs in;
        in.parent_rgb = source;
        in.parent_rgb_bytes = static_cast<size_t>(source_width) * static_cast<size_t>(source_height) * 3;
        in.axes[0] = &ax;
        in.axes[1] = &ay;
        in.out_rgb = output;
        fav::aurora_run(p, in);
    });
}

int fractal_atlas_composite_rgb(const uint8_t* parent, int parent_width, int parent_height, const uint8_t* child,
                                int child_width, int child_height, uint8_t* output, int output_width,
                                int output_height, double parent_zoom, double child_fraction, double child_zoom,
                                int feather, int threads) {
    return atlas_composite_rgb(parent, parent_width, parent_height, child, child_width, child_height, output,
                               output_width, output_height, parent_zoom, child_fraction, child_zoom, feather,
                               threads);
}

int fractal_crop_field(const float* source, int source_width, int source_height, float* output, int output_width,
                       int output_height, double zoom_factor, int threads) {
    return guarded("native raw-field crop failed with an unknown exception", [&] {
        if (!source || !output || !valid_dimensions(source_width, source_height)
            || !valid_dimensions(output_width, output_height) || !valid_threads(threads)
            || !std::isfinite(zoom_factor) || zoom_factor <= 0.0)
            throw std::runtime_error("invalid native raw-field crop dimensions");
        zoom_factor = std::max(zoom_factor, 1.0);
        aurora::Params p = base_params(aurora::kCropField, output_width, output_height);
        p.parent_width = source_width;
        Axis ax, ay;
        fav::atlas::fill_axis(ax, source_width, output_width, zoom_factor);
        fav::atlas::fill_axis(ay, source_height, output_height, zoom_factor);
        AuroraInputs in;
        in.parent_field = source;
        in.parent_field_count = static_cast<size_t>(source_width) * static_cast<size_t>(source_height);
        in.axes[0] = &ax;
        in.axes[1] = &ay;
        in.out_field = output;
        fav::aurora_run(p, in);
    });
}

int fractal_atlas_field(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                        const float* child, int child_width, int child_height, int child_max_iter, float* output,
                        int output_width, int output_height, double parent_zoom, double child_fraction,
                        int palette_max_iter, int threads) {
    // renderer.cpp:18890-18927: child zoom 1, no biases, a 48-pixel seam
    return atlas_field_ex(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                          child_max_iter, output, output_width, output_height, parent_zoom, child_fraction, 1.0, 0.0,
                          0.0, 0.0, palette_max_iter, 48, threads);
}

int fractal_atlas_field_ex(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                           const float* child, int child_width, int child_height, int child_max_iter, float* output,
                           int output_width, int output_height, double parent_zoom, double child_fraction,
                           double child_zoom, double parent_field_bias, double child_field_bias,
                           double output_field_bias, int palette_max_iter, int feather, int threads) {
    return atlas_field_ex(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                          child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                          child_zoom, parent_field_bias, child_field_bias, output_field_bias, palette_max_iter,
                          feather, threads);
}

int fractal_crop_colourise(const float* source, int source_width, int source_height, uint8_t* output,
                           int output_width, int output_height, double zoom_factor, int max_iter, double phase,
                           double vocal, double instrumental, double pitch, int threads) {
    return crop_colourise(source, source_width, source_height, output, output_width, output_height, zoom_factor,
                          max_iter, phase, vocal, instrumental, pitch, threads, nullptr, nullptr);
}

int fractal_crop_colourise_interior(const float* source, int source_width, int source_height, uint8_t* output,
                                    int output_width, int output_height, double zoom_factor, int max_iter,
                                    double phase, double vocal, double instrumental, double pitch, int interior_red,
                                    int interior_green, int interior_blue, int threads) {
    const int interior[3] = {interior_red, interior_green, interior_blue};
    return crop_colourise(source, source_width, source_height, output, output_width, output_height, zoom_factor,
                          max_iter, phase, vocal, instrumental, pitch, threads, interior, nullptr);
}

int fractal_crop_colourise_accents(const float* source, int source_width, int source_height, uint8_t* output,
                                   int output_width, int output_height, double zoom_factor, int max_iter,
                                   double phase, double vocal, double instrumental, double pitch,
                                   const uint8_t* accents, int interior_red, int interior_green, int interior_blue,
                                   int threads) {
    const int interior[3] = {interior_red, interior_green, interior_blue};
    return crop_colourise(source, source_width, source_height, output, output_width, output_height, zoom_factor,
                          max_iter, phase, vocal, instrumental, pitch, threads, interior, accents);
}

int fractal_atlas_colourise(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                            const float* child, int child_width, int child_height, int child_max_iter,
                            uint8_t* output, int output_width, int output_height, double parent_zoom,
                            double child_fraction, int palette_max_iter, double phase, double vocal,
                            double instrumental, double pitch, int threads) {
    return atlas_colourise(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                           child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                           palette_max_iter, phase, vocal, instrumental, pitch, threads, nullptr, nullptr);
}

int fractal_atlas_colourise_interior(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                                     const float* child, int child_width, int child_height, int child_max_iter,
                                     uint8_t* output, int output_width, int output_height, double parent_zoom,
                                     double child_fraction, int palette_max_iter, double phase, double vocal,
                                     double instrumental, double pitch, int interior_red, int interior_green,
                                     int interior_blue, int threads) {
    const int interior[3] = {interior_red, interior_green, interior_blue};
    return atlas_colourise(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                           child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                           palette_max_iter, phase, vocal, instrumental, pitch, threads, interior, nullptr);
}

int fractal_atlas_colourise_accents(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                                    const float* child, int child_width, int child_height, int child_max_iter,
                                    uint8_t* output, int output_width, int output_height, double parent_zoom,
                                    double child_fraction, int palette_max_iter, double phase, double vocal,
                                    double instrumental, double pitch, const uint8_t* accents, int interior_red,
                                    int interior_green, int interior_blue, int threads) {
    if (!accents) {
        fav::set_last_error("ordinary palette accents are required");
        return 1;
    }
    const int interior[3] = {interior_red, interior_green, interior_blue};
    return atlas_colourise(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                           child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                           palette_max_iter, phase, vocal, instrumental, pitch, threads, interior, accents);
}

}  // extern "C"
