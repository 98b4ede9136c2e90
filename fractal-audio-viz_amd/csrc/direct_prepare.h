// direct_prepare.h -- host-side set-up of the shallow direct renderer:
// zoom/centre parsing, viewport geometry and per-row/column coordinates
// (render_direct / render_direct_with_planes, renderer.cpp:10138-10280,
// 10440-10500).  Shared by the C ABI and the CPU harness of the kernel code.
#pragma once

#include <cmath>
#include <locale>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "direct_core.h"

namespace fav {
namespace direct {

// parse_classic_long_double, renderer.cpp:214-226: locale-neutral, the
// whole text must be consumed (trailing white space allowed).
inline bool parse_long_double(const char* text, long double& value) {
    if (!text) return false;
    std::istringstream stream(text);
    stream.imbue(std::locale::classic());
    stream >> value;
    if (stream.fail()) return false;
    stream >> std::ws;
    return stream.eof();
}

// parse_zoom, renderer.cpp:7516-7530.
inline long double parse_zoom_long(const char* text) {
    long double zoom = 0.0L;
    if (!parse_long_double(text, zoom) || !std::isfinite(zoom) || zoom <= 0.0L)
        throw std::runtime_error("invalid Mandelbrot zoom");
    const long double l10 = std::log10(zoom);
    if (!std::isfinite(l10) || l10 < -300.0L || l10 > 9800.0L)
        throw std::runtime_error("Mandelbrot zoom is outside the supported range");
    return zoom;
}

// parse_coordinate, renderer.cpp:228-241.
inline long double parse_coordinate_long(const char* text, const char* label) {
    long double value = 0.0L;
    if (!parse_long_double(text, value) || !std::isfinite(value))
        throw std::runtime_error(std::string("invalid ") + label + " coordinate");
    return value;
}

inline double axis_offset(int index, int dimension, int coordinate_mode) {  // renderer.cpp:142-151
    const double origin = coordinate_mode == 1 ? static_cast<double>(dimension / 2)
                                               : static_cast<double>(dimension - 1) / 2.0;
    return static_cast<double>(index) - origin;
}

struct DirectSetup {
    DirectParams params;
    std::vector<double> x, y;
};

// The per-frame constants and coordinate vectors; pointers in `params` are
// left null for the caller to bind.
inline DirectSetup make_setup_values(int width, int height, long double zoom, long double xl, long double yl,
                                     int max_iter, int formula, double julia_re, double julia_im, double output_bias,
                                     int escape_radius_mode, int coordinate_mode, uint32_t plane_flags) {
    if (escape_radius_mode != 0 && escape_radius_mode != 1) throw std::runtime_error("unknown escape-radius mode");
    if (coordinate_mode != 0 && coordinate_mode != 1) throw std::runtime_error("unknown coordinate mode");
    if (formula < 0 || formula > 3) throw std::runtime_error("unknown fractal formula");
    const double zoom_value = static_cast<double>(zoom);
    const double centre_re = static_cast<double>(xl), centre_im = static_cast<double>(yl);
    if (!std::isfinite(zoom_value) || zoom_value <= 0.0 || !std::isfinite(centre_re) || !std::isfinite(centre_im))
        throw std::runtime_error("direct native coordinates or zoom exceed double range");
    const double height_span = (coordinate_mode == 1 ? 4.0 : 2.8) / zoom_value;
    const double width_span = height_span * static_cast<double>(width) / static_cast<double>(height);
    if (!std::isfinite(height_span) || !std::isfinite(width_span))
        throw std::runtime_error("direct native viewport is outside double range");
    DirectSetup s{};
    s.x.resize(static_cast<size_t>(width));
    s.y.resize(static_cast<size_t>(height));
    for (int px = 0; px < width; ++px)
        s.x[static_cast<size_t>(px)] =
            centre_re + axis_offset(px, width, coordinate_mode) * width_span / static_cast<double>(width);
    for (int py = 0; py < height; ++py)
        s.y[static_cast<size_t>(py)] =
            centre_im + -axis_offset(py, height, coordinate_mode) * height_span / static_cast<double>(height);
    DirectParams& p = s.params;
    p.width = width;
    p.height = height;
    p.max_iter = max_iter;
    p.formula = formula;
    p.plane_flags = plane_flags;
    p.julia_re = julia_re;
    p.julia_im = julia_im;
    p.output_bias = output_bias;
    p.escape2 = escape_r2_double(escape_radius_mode);
    p.safe_escape2 = p.escape2 + 1.0e-7;
    p.log_power = std::log(2.0);
    const FE spacing = fe_from_parts(height_span / static_cast<double>(height), 0);
    p.spacing_m = spacing.m;
    p.spacing_e = spacing.e;
    return s;
}

inline DirectSetup make_setup(int width, int height, const char* zoom_text, const char* x_text, const char* y_text,
                              int max_iter, int formula, double julia_re, double julia_im, double output_bias,
                              int escape_radius_mode, int coordinate_mode, uint32_t plane_flags) {
    const long double zoom = parse_zoom_long(zoom_text);
    const long double xl = parse_coordinate_long(x_text, "real");
    const long double yl = parse_coordinate_long(y_text, "imaginary");
    return make_setup_values(width, height, zoom, xl, yl, max_iter, formula, julia_re, julia_im, output_bias,
                             escape_radius_mode, coordinate_mode, plane_flags);
}

}  // namespace direct
}  // namespace fav
