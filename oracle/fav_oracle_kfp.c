/*
 * fav_oracle_kfp.c -- plain-C restatement of the reference's KFP colour
 * transfer (Kalles SetColor) for a scalar iteration field, and of the
 * 1024-entry cyclic palette LUT.  TEST INFRASTRUCTURE ONLY (fav_oracle.h).
 *
 * Scope: the scalar-field path of fractal_colourise_kfp_impl
 * (renderer.cpp:16314-16545) with planes == NULL, i.e. write_kfp_pixel
 * (renderer.cpp:6696-7289) without orbit planes, phase or texture.
 */
#define _POSIX_C_SOURCE 200809L
#include "fav_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI_D = 3.14159265358979323846;

static void kfp_error(char *error, int capacity, const char *message) {
    if (error && capacity > 0) snprintf(error, (size_t)capacity, "%s", message);
}

static inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline double maxd(double a, double b) { return a < b ? b : a; }  /* std::max */
static inline double mind(double a, double b) { return b < a ? b : a; }  /* std::min */

/* kfp_colour_dither_mask, renderer.cpp:5246-5260 */
static inline float dither_mask(int x, int y, int channel) {
    const unsigned coordinate = ((unsigned)x + (unsigned)channel * 67U + (unsigned)y * 236U) * 119U;
    return (float)(coordinate & 255U) / 256.0F;
}

/* kfp_dithered_srgb_byte, renderer.cpp:5318-5336 */
static inline unsigned char dithered_srgb_byte(float value, int x, int y, int channel) {
    if (!isfinite(value)) value = 0.0F;
    const float mask = dither_mask(x, y, channel);
    const float quantized = 255.0F * value + mask;
    return (unsigned char)clampi((int)floorf(quantized), 0, 255);
}

/* kfp_inside_value, renderer.cpp:5379-5387 */
static inline int inside_value(float value, int max_iter, double bias) {
    return isfinite(value) && (double)value >= (double)max_iter - bias;
}

/* kfp_smooth_offset, renderer.cpp:5400-5428 */
static double smooth_offset_for(const FractalKfpOptions *o) {
    if (o->smooth_method != 0 || !isfinite(o->power) || o->power <= 0.0 || fabs(o->power - 1.0) < 1.0e-12)
        return 0.0;
    double radius = 10000.0;
    switch (o->bailout_radius_preset) {
        case 0: radius = 10000.0; break;
        case 1: radius = 2.0; break;
        case 2: radius = pow(2.0, 1.0 / (o->power - 1.0)); break;
        case 3: radius = o->bailout_radius_custom; break;
        default: return 0.0;
    }
    if (!isfinite(radius) || radius <= 1.0) return 0.0;
    return log(log(radius)) / log(o->power);
}

/* kfp_colour_sample, renderer.cpp:5434-5460 */
static double colour_sample(const float *field, int w, int h, int x, int y, int max_iter, double offset,
                            double bias) {
    x = clampi(x, 0, w - 1);
    y = clampi(y, 0, h - 1);
    const float value = field[(size_t)y * (size_t)w + (size_t)x];
    if (!isfinite(value)) return maxd(0.0, offset);
    if ((double)value >= (double)max_iter - bias) return (double)max_iter + 1.0;
    return maxd(0.0, clampd((double)value + bias, 0.0, (double)max_iter) + offset);
}

/* kfp_difference_magnitude, renderer.cpp:5497-5589 */
static double difference_magnitude(int d, double c, double l, double r, double u, double dn, double tl,
                                   double tr, double bl, double br) {
    const double inverse_sqrt_two = 0.7071067811865475244008443621048490;
    if (d == 0) {
        const double axis_scale = 1.414;
        return fabs(l - c) * axis_scale + fabs(u - c) * axis_scale + fabs(tl - c) + fabs(bl - c);
    }
    if (d == 1) {
        const double sq = (l - c) * (l - c) + (r - c) * (r - c) + (u - c) * (u - c) + (dn - c) * (dn - c)
            + ((tl - c) * (tl - c) + (br - c) * (br - c)) * inverse_sqrt_two * inverse_sqrt_two
            + ((bl - c) * (bl - c) + (tr - c) * (tr - c)) * inverse_sqrt_two * inverse_sqrt_two;
        return sqrt(maxd(0.0, sq * 0.25)) * 2.8284271247461903;
    }
    if (d == 2) {
        const double sq = (r - l) * (r - l) * 0.25 + (dn - u) * (dn - u) * 0.25 + (br - tl) * (br - tl) * 0.125
            + (tr - bl) * (tr - bl) * 0.125;
        return sqrt(maxd(0.0, sq * 0.5)) * 2.8284271247461903;
    }
    if (d == 3) {
        const double sq = (tl - c) * (tl - c) / 2.0 + (l - u) * (l - u) / 2.0;
        return sqrt(maxd(0.0, sq)) * 2.8284271247461903;
    }
    if (d == 4) {
        const double dx = ((u - tl) + (c - l)) * 0.5;
        const double dy = ((l - tl) + (c - u)) * 0.5;
        return sqrt(dx * dx + dy * dy) * 2.8284271247461903;
    }
    if (d == 5) {
        const double dx = (r + tr + br - l - tl - bl) / 6.0;
        const double dy = (dn + bl + br - u - tl - tr) / 6.0;
        return sqrt(dx * dx + dy * dy) * 2.8284271247461903;
    }
    if (d == 6) {
        const double lap = tl + 4.0 * u + tr + 4.0 * l - 20.0 * c + 4.0 * r + bl + 4.0 * dn + br;
        return sqrt(fabs(lap / 6.0 * 1.4426950408889634)) * 2.8284271247461903;
    }
    return 0.0;
}

/* kfp_hsv_to_rgb, renderer.cpp:5591-5644 */
static void hsv_to_rgb(float hue, float saturation, float value, float *r, float *g, float *b) {
    const float scaled_hue = hue * 6.0F;
    const int sector = (int)floorf(scaled_hue);
    float fraction = scaled_hue - (float)sector;
    if ((sector & 1) == 0) fraction = 1.0F - fraction;
    const float minimum = value * (1.0F - saturation);
    const float transition = value * (1.0F - saturation * fraction);
    *r = 0.0F; *g = 0.0F; *b = 0.0F;
    switch (sector) {
        case 0: case 6: *r = minimum; *g = transition; *b = value; break;
        case 1: *r = minimum; *g = value; *b = transition; break;
        case 2: *r = transition; *g = value; *b = minimum; break;
        case 3: *r = value; *g = transition; *b = minimum; break;
        case 4: *r = value; *g = minimum; *b = transition; break;
        case 5: *r = transition; *g = minimum; *b = value; break;
        default: break;
    }
}

typedef struct {
    const float *field;
    const double *samples;
    int w, h, max_iter;
    const FractalKfpOptions *o;
    const unsigned char *lut;
    int lut_size;
    double smooth_offset, tmin, tmax, cosine, sine;
    int needs_difference, needs_slopes;
} kfp_ctx;

static void palette_rgb(const kfp_ctx *k, double input, float *r, float *g, float *b) {
    /* palette_rgb lambda, renderer.cpp:7070-7110 */
    double position = fmod(input, (double)k->lut_size);
    if (position < 0.0) position += (double)k->lut_size;
    const double lower_position = floor(position);
    const int lower = clampi((int)lower_position, 0, k->lut_size - 1);
    const unsigned char *lut = k->lut;
    if (k->o->smooth) {
        double fraction = position - lower_position;
        if (k->o->inverse_transition) fraction = 1.0 - fraction;
        const int upper = (lower + 1) % k->lut_size;
        const double inv = 1.0 - fraction;
        *r = (float)(((double)lut[lower * 3] * inv + (double)lut[upper * 3] * fraction) / 255.0F);
        *g = (float)(((double)lut[lower * 3 + 1] * inv + (double)lut[upper * 3 + 1] * fraction) / 255.0F);
        *b = (float)(((double)lut[lower * 3 + 2] * inv + (double)lut[upper * 3 + 2] * fraction) / 255.0F);
    } else {
        *r = (float)lut[lower * 3] / 255.0F;
        *g = (float)lut[lower * 3 + 1] / 255.0F;
        *b = (float)lut[lower * 3 + 2] / 255.0F;
    }
}

static double reflected(const kfp_ctx *k, int x, int y, int ox, int oy, double centre) {
    const int sx = x + ox, sy = y + oy;
    if (sx >= 0 && sx < k->w && sy >= 0 && sy < k->h) return k->samples[(size_t)sy * (size_t)k->w + (size_t)sx];
    const int px = x - ox, py = y - oy;
    if (px >= 0 && px < k->w && py >= 0 && py < k->h)
        return 2.0 * centre - k->samples[(size_t)py * (size_t)k->w + (size_t)px];
    return centre;
}

/* write_kfp_pixel without planes, renderer.cpp:6696-7289 */
static void kfp_pixel(const kfp_ctx *k, int x, int y, unsigned char *dst) {
    const FractalKfpOptions *o = k->o;
    const size_t idx = (size_t)y * (size_t)k->w + (size_t)x;
    const float raw = k->field[idx];
    int inside;
    double smooth_iter;
    if (k->samples) {
        smooth_iter = k->samples[idx];
        inside = inside_value(raw, k->max_iter, o->field_bias);
    } else {
        inside = inside_value(raw, k->max_iter, o->field_bias);
        const double safe = isfinite(raw) ? clampd((double)raw + o->field_bias, 0.0, (double)k->max_iter) : 0.0;
        smooth_iter = maxd(0.0, safe + k->smooth_offset);
    }
    if (inside) {
        dst[0] = (unsigned char)o->interior_color[0];
        dst[1] = (unsigned char)o->interior_color[1];
        dst[2] = (unsigned char)o->interior_color[2];
        return;
    }
    const double colour_iter = o->flat ? floor(smooth_iter) : smooth_iter;
    const double c = smooth_iter;
    double l = c, r = c, u = c, dn = c, tl = c, tr = c, bl = c, br = c;
    if (k->needs_difference || k->needs_slopes) {
        l = reflected(k, x, y, -1, 0, c);
        u = reflected(k, x, y, 0, -1, c);
        if (k->needs_difference) {
            switch (o->differences) {
                case 0: tl = reflected(k, x, y, -1, -1, c); bl = reflected(k, x, y, -1, 1, c); break;
                case 1: case 2: case 5: case 6:
                    r = reflected(k, x, y, 1, 0, c); dn = reflected(k, x, y, 0, 1, c);
                    tl = reflected(k, x, y, -1, -1, c); tr = reflected(k, x, y, 1, -1, c);
                    bl = reflected(k, x, y, -1, 1, c); br = reflected(k, x, y, 1, 1, c);
                    break;
                case 3: case 4: tl = reflected(k, x, y, -1, -1, c); break;
                default: r = reflected(k, x, y, 1, 0, c); dn = reflected(k, x, y, 0, 1, c); break;
            }
        }
        if (k->needs_slopes) {
            if (x == 0) r = reflected(k, x, y, 1, 0, c);
            if (y == 0) dn = reflected(k, x, y, 0, 1, c);
        }
    }
    double gradient = 0.0;
    if (k->needs_difference) gradient = difference_magnitude(o->differences, c, l, r, u, dn, tl, tr, bl, br);
    double distance = isnan(gradient) ? 0.0 : gradient * (double)k->w / 640.0;
    if (distance < 0.0) distance = 0.0;
    double transfer = colour_iter;
    switch (o->color_method) {
        case 1: transfer = sqrt(maxd(0.0, colour_iter)); break;
        case 2: transfer = pow(maxd(0.0, colour_iter), 1.0 / 3.0); break;
        case 3: transfer = log(maxd(1.0, colour_iter)); break;
        case 4: transfer = 1024.0 * (colour_iter - k->tmin) / maxd(k->tmax - k->tmin, 1.0e-12); break;
        case 5: transfer = mind(distance, 1024.0); break;
        case 6: {
            const double dt = mind(sqrt(maxd(0.0, distance)), 1024.0);
            transfer = dt > o->iter_div ? smooth_iter : dt;
            break;
        }
        case 7: transfer = log(maxd(1.0, distance + 1.0)); break;
        case 8: transfer = sqrt(maxd(0.0, distance)); break;
        case 9: transfer = log(1.0 + log(1.0 + maxd(0.0, colour_iter))); break;
        case 10: transfer = atan(colour_iter); break;
        case 11: transfer = sqrt(sqrt(maxd(0.0, colour_iter))); break;
        default: break;
    }
    if (isnan(transfer) || transfer < 0.0) transfer = 0.0;
    if (o->color_method == 5 || o->color_method == 7 || o->color_method == 8)
        transfer = clampd(transfer, 0.0, 1024.0);
    double palette_value = transfer;
    if (o->iter_div != 1.0) palette_value /= o->iter_div;
    float red, green, blue;
    const double base = palette_value + o->color_offset;
    if (o->multi_color) {
        const double wave_input = o->smooth ? base : floor(base);
        double hs = 0.0, ss = 0.0, vs = 0.0;
        int hc = 0, sc = 0, vc = 0;
        for (unsigned i = 0; i < o->multi_color_count; ++i) {
            const double period = o->multi_color_period[i];
            const double wave = period < 0.0 ? -period / 100.0 : 0.5 + 0.5 * sin(PI_D * wave_input / period);
            switch (o->multi_color_type[i]) {
                case 0: hs += wave; ++hc; break;
                case 1: ss += wave; ++sc; break;
                default: vs += wave; ++vc; break;
            }
        }
        float mr, mg, mb;
        hsv_to_rgb((float)(hc > 0 ? hs / hc : 0.0), (float)(sc > 0 ? ss / sc : 0.0),
                   (float)(vc > 0 ? vs / vc : 0.0), &mr, &mg, &mb);
        if (o->blend_multi_color) {
            palette_rgb(k, base, &red, &green, &blue);
            red = (red + mr) * 0.5F;
            green = (green + mg) * 0.5F;
            blue = (blue + mb) * 0.5F;
        } else {
            red = mr; green = mg; blue = mb;
        }
    } else {
        palette_rgb(k, base, &red, &green, &blue);
    }
    if (k->needs_slopes) {
        const double horizontal = x > 0 ? l - c : c - r;
        const double vertical = y == 0 ? c - dn : u - c;
        const double projected = ((horizontal * k->cosine) + (vertical * k->sine)) * o->slope_power
            * (double)k->w / 640.0;
        const double strength = atan(fabs(projected)) / (PI_D / 2.0) * o->slope_ratio / 100.0;
        if (projected >= 0.0) {
            red = (float)((double)red * (1.0 - strength));
            green = (float)((double)green * (1.0 - strength));
            blue = (float)((double)blue * (1.0 - strength));
        } else {
            red = (float)((double)red * (1.0 - strength) + strength);
            green = (float)((double)green * (1.0 - strength) + strength);
            blue = (float)((double)blue * (1.0 - strength) + strength);
        }
    }
    dst[0] = dithered_srgb_byte(red, x, y, 0);
    dst[1] = dithered_srgb_byte(green, x, y, 1);
    dst[2] = dithered_srgb_byte(blue, x, y, 2);
}

int fo_colourise_kfp(const float *field, uint8_t *output, int width, int height, int max_iter,
                     const FractalKfpOptions *o, const FractalKfpPlanes *planes, const uint8_t *lut,
                     int lut_size, int threads, char *error, int error_capacity) {
    if (!field || !output || !o || !lut || width <= 0 || height <= 0 || lut_size < 2) {
        kfp_error(error, error_capacity, "invalid KFP colour arguments");
        return 1;
    }
    if (planes) {
        kfp_error(error, error_capacity, "the oracle restates the scalar-field KFP path only");
        return 1;
    }
    if (o->field_bias > (double)max_iter) {
        kfp_error(error, error_capacity, "native KFP field bias exceeds iteration cap");
        return 1;
    }
    kfp_ctx k;
    memset(&k, 0, sizeof(k));
    k.field = field;
    k.w = width;
    k.h = height;
    k.max_iter = max_iter;
    k.o = o;
    k.lut = lut;
    k.lut_size = lut_size;
    k.smooth_offset = smooth_offset_for(o);
    k.cosine = 1.0;
    k.sine = 0.0;
    if (o->slopes) {  /* kfp_slope_direction, renderer.cpp:7509-7514 */
        const double angle = o->slope_angle * 3.14159265358979323846 / 180.0;
        k.cosine = cos(angle);
        k.sine = sin(angle);
    }
    k.needs_difference = o->color_method >= 5 && o->color_method <= 8;
    k.needs_slopes = o->slopes && o->slope_power > 0.0 && o->slope_ratio > 0.0;
#ifdef _OPENMP
    if (threads > 0) { omp_set_dynamic(0); omp_set_num_threads(threads); }
#else
    (void)threads;
#endif
    const long long n = (long long)width * height;
    double *samples = NULL;
    if (k.needs_difference || k.needs_slopes) {
        samples = (double *)malloc(sizeof(double) * (size_t)n);
        if (!samples) { kfp_error(error, error_capacity, "out of memory"); return 1; }
#pragma omp parallel for schedule(static)
        for (long long p = 0; p < n; ++p)
            samples[p] = colour_sample(field, width, height, (int)(p % width), (int)(p / width), max_iter,
                                       k.smooth_offset, o->field_bias);
        k.samples = samples;
    }
    k.tmin = 0.0;
    k.tmax = 1.0;
    if (o->color_method == 4) {  /* kfp_transfer_bounds, renderer.cpp:7296-7334 */
        double mn = INFINITY, mx = -INFINITY;
        for (long long p = 0; p < n; ++p) {
            const float raw = field[p];
            if (inside_value(raw, max_iter, o->field_bias)) continue;
            const double safe = isfinite(raw) ? clampd((double)raw + o->field_bias, 0.0, (double)max_iter) : 0.0;
            const double value = floor(safe + k.smooth_offset);
            if (value >= (double)(max_iter - 1)) continue;
            mn = mind(mn, value);
            mx = maxd(mx, value);
        }
        if (isfinite(mn) && isfinite(mx)) { k.tmin = mn; k.tmax = mx; }
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) kfp_pixel(&k, x, y, output + ((size_t)y * (size_t)width + (size_t)x) * 3U);
    free(samples);
    return 0;
}

/* _kfp_palette_lut, visualizer.py:13492-13521: Kalles' cyclic sine
 * interpolation of the key colours, truncated to bytes. */
int fo_kfp_palette_lut(const uint8_t *stops, int stop_count, uint8_t *lut, int lut_size) {
    if (!stops || !lut || stop_count < 2 || lut_size < 2) return 1;
    for (int i = 0; i < lut_size; ++i) {
        const double position = (double)i * (double)stop_count / (double)lut_size;
        const int left = (int)floor(position);
        const int right = (left + 1) % stop_count;
        double fraction = position - (double)left;
        fraction = sin((fraction - 0.5) * PI_D) / 2.0 + 0.5;
        for (int ch = 0; ch < 3; ++ch) {
            double v = (double)stops[left * 3 + ch] * (1.0 - fraction) + (double)stops[right * 3 + ch] * fraction;
            v = clampd(v, 0.0, 255.0);
            lut[i * 3 + ch] = (uint8_t)v;
        }
    }
    return 0;
}


/* ---- raw KFP atlas colour pass ------------------------------------------ */

/* fill_bilinear_axis_window, renderer.cpp:4969-5002 (fill_bilinear_axis,
 * renderer.cpp:4931-4962, is the window with origin 0 and span equal to the
 * destination size). */
static void atlas_axis(int *index0, int *index1, float *weight, int source_size, int destination_size,
                       double origin, double span, double zoom) {
    zoom = zoom > 1.0 ? zoom : 1.0;
    const double crop = (double)source_size / zoom;
    const double left = ((double)source_size - crop) * 0.5;
    for (int d = 0; d < destination_size; ++d) {
        const double logical = origin + (double)d + 0.5;
        double source = clampd(left + logical * crop / span - 0.5, 0.0, (double)(source_size - 1));
        const int i0 = (int)floor(source);
        index0[d] = i0;
        index1[d] = i0 + 1 < source_size - 1 ? i0 + 1 : source_size - 1;
        weight[d] = (float)(source - (double)i0);
    }
}

/* fill_bilinear_axis, renderer.cpp:4931-4962: the plain centred crop. */
static void atlas_axis_plain(int *index0, int *index1, float *weight, int source_size, int destination_size,
                             double zoom) {
    zoom = zoom > 1.0 ? zoom : 1.0;
    const double inverse_zoom = 1.0 / zoom;
    const double crop = (double)source_size * inverse_zoom;
    const double left = ((double)source_size - crop) * 0.5;
    const double step = crop / (double)destination_size;
    for (int d = 0; d < destination_size; ++d) {
        double source = clampd(left + ((double)d + 0.5) * step - 0.5, 0.0, (double)(source_size - 1));
        const int i0 = (int)floor(source);
        index0[d] = i0;
        index1[d] = i0 + 1 < source_size - 1 ? i0 + 1 : source_size - 1;
        weight[d] = (float)(source - (double)i0);
    }
}

/* sample_bilinear_mapped_preserving_interior, renderer.cpp:5146-5208. */
static float atlas_sample(const float *src, int width, const int *x0, const int *x1, const float *xw,
                          const int *y0, const int *y1, const float *yw, int x, int y, int max_iter, int *inside,
                          double bias) {
    const double wx = (double)xw[x], wy = (double)yw[y];
    const float *top = src + (size_t)y0[y] * (size_t)width;
    const float *bottom = src + (size_t)y1[y] * (size_t)width;
    const float values[4] = {top[x0[x]], top[x1[x]], bottom[x0[x]], bottom[x1[x]]};
    const double weights[4] = {(1.0 - wy) * (1.0 - wx), (1.0 - wy) * wx, wy * (1.0 - wx), wy * wx};
    double interior = 0.0, exterior = 0.0, value = 0.0;
    for (int k = 0; k < 4; ++k) {
        const float v = values[k];
        if (isfinite(v) && (double)v >= (double)max_iter - bias) {
            interior += weights[k];
        } else if (isfinite(v)) {
            exterior += weights[k];
            value += ((double)v + bias) * weights[k];
        }
    }
    if (interior >= 0.5) {
        *inside = 1;
        return (float)max_iter;
    }
    *inside = 0;
    if (exterior <= 1.0e-12) return 0.0f;
    return (float)(value / exterior);
}

/* encode_render_value, renderer.cpp:119-125 (long double subtraction). */
static float atlas_encode(long double value, double bias) { return (float)(value - (long double)bias); }

/* kfp_atlas_seam_feather, renderer.cpp:99-110 (maximum 64, divisor 2). */
static int atlas_feather(int width, int height) {
    const int minimum = width < height ? width : height;
    if (minimum <= 0) return 0;
    int requested = minimum / 2 < 64 ? minimum / 2 : 64;
    if (requested < 2) requested = 2;
    const int plateau = (minimum - 1) / 2 > 0 ? (minimum - 1) / 2 : 0;
    return requested < plateau ? requested : plateau;
}

int fo_atlas_colourise_kfp_raw(const float *parent, int parent_width, int parent_height, int parent_max_iter,
                               const float *child, int child_width, int child_height, int child_max_iter,
                               uint8_t *output, int output_width, int output_height, double parent_zoom,
                               double child_fraction, double child_zoom, int max_iter,
                               const FractalKfpOptions *o, const uint8_t *lut, int lut_size, int threads,
                               char *error, int error_capacity) {
    /* fractal_atlas_colourise_kfp_raw, renderer.cpp:17969-18265 */
    if (!parent || !output || !o || !lut || parent_width <= 0 || parent_height <= 0 || output_width <= 0
        || output_height <= 0 || !isfinite(parent_zoom) || parent_zoom <= 0.0 || !isfinite(child_fraction)
        || child_fraction < 0.0 || child_fraction > 1.0 || !isfinite(child_zoom) || child_zoom <= 0.0) {
        kfp_error(error, error_capacity, "invalid native raw KFP atlas dimensions or controls");
        return 1;
    }
    parent_zoom = parent_zoom > 1.0 ? parent_zoom : 1.0;
    child_zoom = child_zoom > 1.0 ? child_zoom : 1.0;
    const int use_child = child != NULL && child_fraction > 0.0;
    if (use_child && (child_width <= 0 || child_height <= 0 || child_max_iter <= 0)) {
        kfp_error(error, error_capacity, "invalid native raw KFP atlas child tile");
        return 1;
    }
    if (o->field_bias > (double)max_iter) {
        kfp_error(error, error_capacity, "native KFP field bias exceeds iteration cap");
        return 1;
    }
    const int centered = o->field_bias > 0.0;
    const double parent_bias = centered ? (double)parent_max_iter : 0.0;
    const double child_bias = centered ? (double)child_max_iter : 0.0;
    int visible_w = 0, visible_h = 0, visible_left = 0, visible_top = 0, halo = 0;
    if (use_child) {
        visible_w = (int)lround((double)output_width * child_fraction);
        visible_h = (int)lround((double)output_height * child_fraction);
        if (visible_w < 1) visible_w = 1;
        if (visible_h < 1) visible_h = 1;
        visible_left = (output_width - visible_w) / 2;
        visible_top = (output_height - visible_h) / 2;
        int m = visible_left;
        if (output_width - visible_left - visible_w < m) m = output_width - visible_left - visible_w;
        if (visible_top < m) m = visible_top;
        if (output_height - visible_top - visible_h < m) m = output_height - visible_top - visible_h;
        halo = m < 2 ? m : 2;
    }
    const int child_left = visible_left - halo, child_top = visible_top - halo;
    const int sampled_w = use_child ? visible_w + 2 * halo : 0, sampled_h = use_child ? visible_h + 2 * halo : 0;
    const int feather = use_child ? atlas_feather(visible_w, visible_h) : 0;

    const size_t n = (size_t)output_width * (size_t)output_height;
    int *px0 = malloc(sizeof(int) * (size_t)output_width), *px1 = malloc(sizeof(int) * (size_t)output_width);
    int *py0 = malloc(sizeof(int) * (size_t)output_height), *py1 = malloc(sizeof(int) * (size_t)output_height);
    float *pxw = malloc(sizeof(float) * (size_t)output_width), *pyw = malloc(sizeof(float) * (size_t)output_height);
    float *view = malloc(sizeof(float) * n);
    int sw = sampled_w > 0 ? sampled_w : 1, sh = sampled_h > 0 ? sampled_h : 1;
    int *cx0 = malloc(sizeof(int) * (size_t)sw), *cx1 = malloc(sizeof(int) * (size_t)sw);
    int *cy0 = malloc(sizeof(int) * (size_t)sh), *cy1 = malloc(sizeof(int) * (size_t)sh);
    float *cxw = malloc(sizeof(float) * (size_t)sw), *cyw = malloc(sizeof(float) * (size_t)sh);
    float *child_view = malloc(sizeof(float) * (size_t)sw * (size_t)sh);
    int status = 1;
    if (!px0 || !px1 || !py0 || !py1 || !pxw || !pyw || !view || !cx0 || !cx1 || !cy0 || !cy1 || !cxw || !cyw
        || !child_view) {
        kfp_error(error, error_capacity, "out of memory");
        goto done;
    }
    atlas_axis_plain(px0, px1, pxw, parent_width, output_width, parent_zoom);
    atlas_axis_plain(py0, py1, pyw, parent_height, output_height, parent_zoom);
    for (int y = 0; y < output_height; ++y)
        for (int x = 0; x < output_width; ++x) {
            int inside = 0;
            const float smooth = atlas_sample(parent, parent_width, px0, px1, pxw, py0, py1, pyw, x, y,
                                              parent_max_iter, &inside, parent_bias);
            view[(size_t)y * (size_t)output_width + (size_t)x] = inside
                ? atlas_encode((long double)max_iter, o->field_bias)
                : atlas_encode((long double)smooth, o->field_bias);
        }
    if (use_child) {
        atlas_axis(cx0, cx1, cxw, child_width, sampled_w, -(double)halo, (double)visible_w, child_zoom);
        atlas_axis(cy0, cy1, cyw, child_height, sampled_h, -(double)halo, (double)visible_h, child_zoom);
        for (int y = 0; y < sampled_h; ++y)
            for (int x = 0; x < sampled_w; ++x) {
                int inside = 0;
                const float smooth = atlas_sample(child, child_width, cx0, cx1, cxw, cy0, cy1, cyw, x, y,
                                                  child_max_iter, &inside, child_bias);
                child_view[(size_t)y * (size_t)sampled_w + (size_t)x] = inside
                    ? atlas_encode((long double)max_iter, o->field_bias)
                    : atlas_encode((long double)smooth, o->field_bias);
            }
        for (int y = visible_top; y < visible_top + visible_h; ++y)
            for (int x = visible_left; x < visible_left + visible_w; ++x) {
                const size_t at = (size_t)y * (size_t)output_width + (size_t)x;
                const float cv = child_view[(size_t)(y - child_top) * (size_t)sampled_w + (size_t)(x - child_left)];
                float alpha = 1.0f;
                if (feather >= 2) {
                    const int vx = x - visible_left, vy = y - visible_top;
                    int edge = vx < visible_w - 1 - vx ? vx : visible_w - 1 - vx;
                    const int ey = vy < visible_h - 1 - vy ? vy : visible_h - 1 - vy;
                    if (ey < edge) edge = ey;
                    float linear = (float)edge / (float)feather;
                    linear = linear < 0.0f ? 0.0f : (linear > 1.0f ? 1.0f : linear);
                    alpha = linear * linear * (3.0f - 2.0f * linear);
                }
                if (alpha >= 0.999999f) view[at] = cv;
                else view[at] = view[at] * (1.0f - alpha) + cv * alpha;
            }
    }
    status = fo_colourise_kfp(view, output, output_width, output_height, max_iter, o, NULL, lut, lut_size, threads,
                              error, error_capacity);
done:
    free(px0); free(px1); free(py0); free(py1); free(pxw); free(pyw); free(view);
    free(cx0); free(cx1); free(cy0); free(cy1); free(cxw); free(cyw); free(child_view);
    return status;
}
