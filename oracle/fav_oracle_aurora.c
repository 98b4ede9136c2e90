/* fav_oracle_aurora.c -- CPU restatement of the reference's ordinary
 * (Aurora) palette and per-frame crop / atlas compositors, the checker of
 * the HIP compositor kernels.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product links, loads or calls this file.
 *
 * Every fo_* function takes the arguments of the renderer.h entry point of
 * the same name without the fo_ prefix (renderer.h:400-449, 713-835,
 * 943-1068) and returns 0, or 1 for invalid arguments (no error text).
 * Each follows the reference file:line cited beside it, operation by
 * operation, in plain C with the host libm. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FO_MAX_ITER 10000000
#define FO_MAX_PIXELS 100000000ULL

static int ok_dims(int w, int h) { /* renderer.cpp:170-175 */
    return w > 0 && h > 0 && (uint64_t)w <= FO_MAX_PIXELS / (uint64_t)h;
}
static int ok_iter(int n) { return n > 0 && n <= FO_MAX_ITER; }
static int ok_controls(double a, double b, double c, double d) {
    return isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d);
}
static int ok_interior(const int *c) {
    return !c || (c[0] >= 0 && c[0] <= 255 && c[1] >= 0 && c[1] <= 255 && c[2] >= 0 && c[2] <= 255);
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int mini(int a, int b) { return a < b ? a : b; }
static int maxi(int a, int b) { return a > b ? a : b; }

/* colour_byte and rounded_colour_byte, renderer.cpp:4807-4809, 5237-5242 */
static uint8_t byte_trunc(double v) { return (uint8_t)clampd(v, 0.0, 255.0); }
static uint8_t byte_round(double v) {
    int i = (int)nearbyint(clampd(v, 0.0, 255.0));
    return (uint8_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}

/* pitch_hue_angle and rotate_hue_rgb, renderer.cpp:4813-4839 */
static double hue_angle(double pitch) {
    return clampd(2.0 * (pitch - 0.5), -1.0, 1.0) * 0.16 * 6.283185307179586476925286766559005768;
}
static void hue_rotate(double *rgb, double c, double s) {
    const double y = 0.299 * rgb[0] + 0.587 * rgb[1] + 0.114 * rgb[2];
    const double i = 0.596 * rgb[0] - 0.275 * rgb[1] - 0.321 * rgb[2];
    const double q = 0.212 * rgb[0] - 0.523 * rgb[1] + 0.311 * rgb[2];
    const double ri = i * c - q * s, rq = i * s + q * c;
    rgb[0] = y + 0.956 * ri + 0.621 * rq;
    rgb[1] = y - 0.272 * ri - 0.647 * rq;
    rgb[2] = y - 1.106 * ri + 1.703 * rq;
}

/* aurora_palette_for with colour_basis_for, renderer.cpp:4788-4805, 4841-4923.
 * Returns a malloc'd size*3 table. */
static uint8_t *palette_new(int max_iter, double phase, double vocal, double pitch, const uint8_t *accents,
                            int *size_out) {
    int size = max_iter * 4;
    size = size < 4096 ? 4096 : (size > 65536 ? 65536 : size);
    uint8_t *t = (uint8_t *)malloc((size_t)size * 3);
    if (!t) return NULL;
    const double den = (double)(size - 1 > 1 ? size - 1 : 1);
    const double vm = clampd(vocal, 0.0, 1.0), split = 5.0 * vm * vm;
    const double rc = cos(phase), rs = sin(phase);
    const double gc = cos(phase + split * 0.4), gs = sin(phase + split * 0.4);
    const double bc = cos(phase + split), bs = sin(phase + split);
    const double angle_h = hue_angle(pitch);
    const int rotate = fabs(angle_h) > 1.0e-15;
    const double hc = rotate ? cos(angle_h) : 1.0, hs = rotate ? sin(angle_h) : 0.0;
    for (int k = 0; k < size; ++k) {
        const double angle = 0.15 * (double)max_iter * (double)k / den;
        const double c = (double)(float)cos(angle), s = (double)(float)sin(angle);
        double v[3] = {(0.5 - 0.5 * (c * rc + s * rs)) * 140.0, (0.5 - 0.5 * (c * gc + s * gs)) * 140.0,
                       (0.5 - 0.5 * (c * bc + s * bs)) * 140.0};
        if (accents) {
            const double w0 = clampd((double)byte_trunc(v[0]) / 140.0, 0.0, 1.0);
            const double w1 = clampd((double)byte_trunc(v[1]) / 140.0, 0.0, 1.0);
            const double w2 = clampd((double)byte_trunc(v[2]) / 140.0, 0.0, 1.0);
            double o[3];
            for (int ch = 0; ch < 3; ++ch)
                o[ch] = (w0 * (double)accents[ch] + w1 * (double)accents[3 + ch] + w2 * (double)accents[6 + ch]) / 1.8;
            memcpy(v, o, sizeof(v));
        }
        if (rotate) hue_rotate(v, hc, hs);
        for (int ch = 0; ch < 3; ++ch) t[(size_t)k * 3 + (size_t)ch] = byte_trunc(v[ch]);
    }
    *size_out = size;
    return t;
}

/* fill_bilinear_axis, renderer.cpp:4931-4962 */
typedef struct { int *i0, *i1; float *w; } axis_t;
static int axis_new(axis_t *a, int src, int dst, double zoom) {
    a->i0 = (int *)malloc(sizeof(int) * (size_t)dst);
    a->i1 = (int *)malloc(sizeof(int) * (size_t)dst);
    a->w = (float *)malloc(sizeof(float) * (size_t)dst);
    if (!a->i0 || !a->i1 || !a->w) return 0;
    zoom = zoom > 1.0 ? zoom : 1.0;
    const double crop = (double)src * (1.0 / zoom);
    const double left = ((double)src - crop) * 0.5;
    const double step = crop / (double)dst;
    for (int d = 0; d < dst; ++d) {
        double s = left + ((double)d + 0.5) * step - 0.5;
        s = s < 0.0 ? 0.0 : (s > (double)(src - 1) ? (double)src - 1.0 : s);
        const int i = (int)floor(s);
        a->i0[d] = i;
        a->i1[d] = i + 1 < src - 1 ? i + 1 : src - 1;
        a->w[d] = (float)(s - (double)i);
    }
    return 1;
}
static void axis_free(axis_t *a) { free(a->i0); free(a->i1); free(a->w); memset(a, 0, sizeof(*a)); }

/* sample_bilinear_mapped, renderer.cpp:5053-5074 */
static float sample_plain(const float *src, int w, const axis_t *ax, const axis_t *ay, int x, int y) {
    const float *t = src + (size_t)ay->i0[y] * (size_t)w, *b = src + (size_t)ay->i1[y] * (size_t)w;
    const float wx = ax->w[x], wy = ay->w[y];
    const float tv = t[ax->i0[x]] * (1.0F - wx) + t[ax->i1[x]] * wx;
    const float bv = b[ax->i0[x]] * (1.0F - wx) + b[ax->i1[x]] * wx;
    return tv * (1.0F - wy) + bv * wy;
}

/* sample_bilinear_mapped_preserving_interior, renderer.cpp:5146-5208 */
static float sample_interior(const float *src, int w, const axis_t *ax, const axis_t *ay, int x, int y, int cap,
                             int *inside, double bias) {
    const float *t = src + (size_t)ay->i0[y] * (size_t)w, *b = src + (size_t)ay->i1[y] * (size_t)w;
    const double wx = (double)ax->w[x], wy = (double)ay->w[y];
    const float v[4] = {t[ax->i0[x]], t[ax->i1[x]], b[ax->i0[x]], b[ax->i1[x]]};
    const double wt[4] = {(1.0 - wy) * (1.0 - wx), (1.0 - wy) * wx, wy * (1.0 - wx), wy * wx};
    double in_w = 0.0, ex_w = 0.0, ex_v = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (isfinite(v[k]) && (double)v[k] >= (double)cap - bias)
            in_w += wt[k];
        else if (isfinite(v[k])) {
            ex_w += wt[k];
            ex_v += ((double)v[k] + bias) * wt[k];
        }
    }
    if (in_w >= 0.5) { *inside = 1; return (float)cap; }
    *inside = 0;
    if (ex_w <= 1.0e-12) return 0.0F;
    return (float)(ex_v / ex_w);
}

/* sample_rgb_bilinear_mapped, renderer.cpp:5262-5300 */
static void sample_rgb(const uint8_t *src, int w, const axis_t *ax, const axis_t *ay, int x, int y, uint8_t *d) {
    const size_t W = (size_t)w;
    const uint8_t *tl = src + ((size_t)ay->i0[y] * W + (size_t)ax->i0[x]) * 3;
    const uint8_t *tr = src + ((size_t)ay->i0[y] * W + (size_t)ax->i1[x]) * 3;
    const uint8_t *bl = src + ((size_t)ay->i1[y] * W + (size_t)ax->i0[x]) * 3;
    const uint8_t *br = src + ((size_t)ay->i1[y] * W + (size_t)ax->i1[x]) * 3;
    const float wx = ax->w[x], wy = ay->w[y], ix = 1.0F - wx, iy = 1.0F - wy;
    for (int c = 0; c < 3; ++c) {
        const float t = (float)tl[c] * ix + (float)tr[c] * wx;
        const float b = (float)bl[c] * ix + (float)br[c] * wx;
        d[c] = byte_round((double)(t * iy + b * wy));
    }
}

/* write_colour_pixel, renderer.cpp:5210-5235 */
static void colour(float smooth, int cap, const uint8_t *pal, int size, float scale, const int *interior, uint8_t *d) {
    if (!isfinite(smooth) || smooth >= (float)cap) {
        for (int c = 0; c < 3; ++c) d[c] = interior ? (uint8_t)interior[c] : 0;
        return;
    }
    const double s = (double)smooth * (double)scale;
    const int k = s >= (double)(size - 1) ? size - 1 : (!isfinite(s) || s <= 0.0 ? 0 : (int)s);
    memcpy(d, pal + (size_t)k * 3, 3);
}

/* colourise_field_impl, renderer.cpp:13771-13825 */
static int colourise_impl(const float *field, uint8_t *out, int w, int h, int max_iter, double phase, double vocal,
                          double instrumental, double pitch, int threads, const int *interior,
                          const uint8_t *accents) {
    if (!field || !out || !ok_dims(w, h) || !ok_iter(max_iter) || threads < 0 || threads > 4096
        || !ok_controls(phase, vocal, instrumental, pitch) || !ok_interior(interior))
        return 1;
    int size = 0;
    uint8_t *pal = palette_new(max_iter, phase, vocal, pitch, accents, &size);
    if (!pal) return 1;
    const float scale = (float)(size - 1) / (float)max_iter;
    for (size_t i = 0; i < (size_t)w * (size_t)h; ++i) colour(field[i], max_iter, pal, size, scale, interior, out + i * 3);
    free(pal);
    return 0;
}

int fo_colourise(const float *field, uint8_t *out, int w, int h, int max_iter, double phase, double vocal,
                 double instrumental, double pitch, int threads) {
    return colourise_impl(field, out, w, h, max_iter, phase, vocal, instrumental, pitch, threads, NULL, NULL);
}

int fo_colourise_accents(const float *field, uint8_t *out, int w, int h, int max_iter, double phase, double vocal,
                         double instrumental, double pitch, const uint8_t *accents, int ir, int ig, int ib,
                         int threads) { /* renderer.cpp:16210-16248 */
    if (!accents) return 1;
    const int interior[3] = {ir, ig, ib};
    return colourise_impl(field, out, w, h, max_iter, phase, vocal, instrumental, pitch, threads, interior, accents);
}

int fo_apply_aurora_accents(uint8_t *out, int w, int h, const uint8_t *accents, double pitch, int threads) {
    /* renderer.cpp:16250-16300 */
    if (!out || !accents || !ok_dims(w, h) || threads < 0 || threads > 4096 || !isfinite(pitch)) return 1;
    const double a = hue_angle(pitch);
    const int rotate = fabs(a) > 1.0e-15;
    const double hc = rotate ? cos(a) : 1.0, hs = rotate ? sin(a) : 0.0;
    for (size_t i = 0; i < (size_t)w * (size_t)h; ++i) {
        uint8_t *d = out + i * 3;
        const double w0 = clampd((double)d[0] / 140.0, 0.0, 1.0), w1 = clampd((double)d[1] / 140.0, 0.0, 1.0);
        const double w2 = clampd((double)d[2] / 140.0, 0.0, 1.0);
        double v[3];
        for (int c = 0; c < 3; ++c)
            v[c] = (w0 * (double)accents[c] + w1 * (double)accents[3 + c] + w2 * (double)accents[6 + c]) / 1.8;
        if (rotate) hue_rotate(v, hc, hs);
        for (int c = 0; c < 3; ++c) d[c] = byte_round(v[c]);
    }
    return 0;
}

int fo_crop_rgb(const uint8_t *src, int sw, int sh, uint8_t *out, int ow, int oh, double zoom, int threads) {
    /* renderer.cpp:18399-18465 */
    if (!src || !out || !ok_dims(sw, sh) || !ok_dims(ow, oh) || threads < 0 || threads > 4096 || !isfinite(zoom)
        || zoom <= 0.0)
        return 1;
    zoom = zoom > 1.0 ? zoom : 1.0;
    if (sw == ow && sh == oh && zoom == 1.0) {
        memcpy(out, src, (size_t)ow * (size_t)oh * 3);
        return 0;
    }
    axis_t ax = {0}, ay = {0};
    int ok = axis_new(&ax, sw, ow, zoom) && axis_new(&ay, sh, oh, zoom);
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) sample_rgb(src, sw, &ax, &ay, x, y, out + ((size_t)y * (size_t)ow + (size_t)x) * 3);
    axis_free(&ax);
    axis_free(&ay);
    return ok ? 0 : 1;
}

int fo_atlas_composite_rgb(const uint8_t *parent, int pw, int ph, const uint8_t *child, int cw, int ch, uint8_t *out,
                           int ow, int oh, double parent_zoom, double fraction, double child_zoom, int feather,
                           int threads) { /* renderer.cpp:18467-18627 */
    if (!parent || !out || !ok_dims(pw, ph) || !ok_dims(ow, oh) || threads < 0 || threads > 4096
        || !isfinite(parent_zoom) || parent_zoom <= 0.0 || !isfinite(fraction) || fraction < 0.0 || fraction > 1.0
        || !isfinite(child_zoom) || child_zoom <= 0.0 || feather < 0)
        return 1;
    parent_zoom = parent_zoom > 1.0 ? parent_zoom : 1.0;
    child_zoom = child_zoom > 1.0 ? child_zoom : 1.0;
    const int use = child != NULL && fraction > 0.0;
    if ((!use && fraction > 0.0) || (!use && child) || (!use && (cw || ch))) return 1;
    if (use && (!ok_dims(cw, ch) || feather > mini(ow, oh))) return 1;
    const int vw = use ? maxi(1, (int)lround((double)ow * fraction)) : 0;
    const int vh = use ? maxi(1, (int)lround((double)oh * fraction)) : 0;
    const int cl = use ? (ow - vw) / 2 : 0, ct = use ? (oh - vh) / 2 : 0;
    const int fe = use ? mini(feather, mini(vw, vh)) : 0;
    axis_t px = {0}, py = {0}, cx = {0}, cy = {0};
    int ok = axis_new(&px, pw, ow, parent_zoom) && axis_new(&py, ph, oh, parent_zoom);
    if (ok && use) ok = axis_new(&cx, cw, vw, child_zoom) && axis_new(&cy, ch, vh, child_zoom);
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            uint8_t *d = out + ((size_t)y * (size_t)ow + (size_t)x) * 3;
            sample_rgb(parent, pw, &px, &py, x, y, d);
            if (!use || x < cl || x >= cl + vw || y < ct || y >= ct + vh) continue;
            uint8_t c3[3];
            sample_rgb(child, cw, &cx, &cy, x - cl, y - ct, c3);
            double alpha = 1.0;
            if (fe >= 2) {
                const int e = mini(mini(x - cl, vw - 1 - (x - cl)), mini(y - ct, vh - 1 - (y - ct)));
                const double l = clampd((double)e / (double)fe, 0.0, 1.0);
                alpha = l * l * (3.0 - 2.0 * l);
            }
            for (int c = 0; c < 3; ++c) d[c] = byte_round((double)d[c] * (1.0 - alpha) + (double)c3[c] * alpha);
        }
    axis_free(&px); axis_free(&py); axis_free(&cx); axis_free(&cy);
    return ok ? 0 : 1;
}

int fo_crop_field(const float *src, int sw, int sh, float *out, int ow, int oh, double zoom, int threads) {
    /* renderer.cpp:18631-18689 */
    if (!src || !out || !ok_dims(sw, sh) || !ok_dims(ow, oh) || threads < 0 || threads > 4096 || !isfinite(zoom)
        || zoom <= 0.0)
        return 1;
    zoom = zoom > 1.0 ? zoom : 1.0;
    axis_t ax = {0}, ay = {0};
    int ok = axis_new(&ax, sw, ow, zoom) && axis_new(&ay, sh, oh, zoom);
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) out[(size_t)y * (size_t)ow + (size_t)x] = sample_plain(src, sw, &ax, &ay, x, y);
    axis_free(&ax);
    axis_free(&ay);
    return ok ? 0 : 1;
}

/* separable smoothstep seam weights in float, renderer.cpp:18791-18810 */
static float *seam(int n, int feather) {
    float *v = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; v && i < n; ++i) {
        const int e = mini(i, n - 1 - i);
        float l = (float)e / (float)feather;
        l = l < 1.0F ? l : 1.0F;
        v[i] = l * l * (3.0F - 2.0F * l);
    }
    return v;
}

int fo_atlas_field_ex(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch, int cmax,
                      float *out, int ow, int oh, double parent_zoom, double fraction, double child_zoom, double pbias,
                      double cbias, double obias, int palette_max_iter, int feather, int threads) {
    /* renderer.cpp:18697-18888 */
    if (!parent || !out || !ok_dims(pw, ph) || !ok_dims(ow, oh) || !ok_iter(pmax) || threads < 0 || threads > 4096
        || !isfinite(parent_zoom) || parent_zoom <= 0.0 || !isfinite(fraction) || fraction < 0.0 || fraction > 1.0
        || !isfinite(child_zoom) || child_zoom <= 0.0 || !isfinite(pbias) || pbias < 0.0 || pbias > (double)pmax
        || !isfinite(cbias) || cbias < 0.0 || !isfinite(obias) || obias < 0.0 || feather < 0
        || !ok_iter(palette_max_iter))
        return 1;
    const int use = child != NULL && fraction > 0.0;
    if (use && (!ok_dims(cw, ch) || !ok_iter(cmax) || cbias > (double)cmax)) return 1;
    const int eff = maxi(palette_max_iter, maxi(pmax, use ? cmax : 0));
    if (!ok_iter(eff) || obias > (double)eff) return 1;
    const int full = use && fraction >= 0.999999;
    const int vw = use ? maxi(1, (int)lround((double)ow * fraction)) : 0;
    const int vh = use ? maxi(1, (int)lround((double)oh * fraction)) : 0;
    const int cl = use ? (ow - vw) / 2 : 0, ct = use ? (oh - vh) / 2 : 0;
    const int fe = use && !full ? mini(feather, mini(vw / 2, vh / 2)) : 0;
    axis_t px = {0}, py = {0}, cx = {0}, cy = {0};
    int ok = 1;
    if (!full) ok = axis_new(&px, pw, ow, parent_zoom) && axis_new(&py, ph, oh, parent_zoom);
    const double cz = child_zoom > 1.0 ? child_zoom : 1.0;
    if (ok && use) ok = axis_new(&cx, cw, full ? ow : vw, cz) && axis_new(&cy, ch, full ? oh : vh, cz);
    float *ex = NULL, *ey = NULL;
    if (ok && use && !full && fe >= 2) { ex = seam(vw, fe); ey = seam(vh, fe); ok = ex && ey; }
    const float inner = (float)((double)eff - obias), fb = (float)obias;
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            float v;
            int inside = 0;
            if (full) {
                const float s = sample_interior(child, cw, &cx, &cy, x, y, cmax, &inside, cbias);
                v = inside ? inner : s - fb;
            } else {
                const float s = sample_interior(parent, pw, &px, &py, x, y, pmax, &inside, pbias);
                v = inside ? inner : s - fb;
                if (use && x >= cl && x < cl + vw && y >= ct && y < ct + vh) {
                    int ci = 0;
                    const float cs = sample_interior(child, cw, &cx, &cy, x - cl, y - ct, cmax, &ci, cbias);
                    const float cv = ci ? inner : cs - fb;
                    float a = 1.0F;
                    if (fe >= 2) a = ex[x - cl] < ey[y - ct] ? ex[x - cl] : ey[y - ct];
                    v = v * (1.0F - a) + cv * a;
                }
            }
            out[(size_t)y * (size_t)ow + (size_t)x] = v;
        }
    free(ex); free(ey);
    axis_free(&px); axis_free(&py); axis_free(&cx); axis_free(&cy);
    return ok ? 0 : 1;
}

int fo_atlas_field(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch, int cmax,
                   float *out, int ow, int oh, double parent_zoom, double fraction, int palette_max_iter,
                   int threads) { /* renderer.cpp:18890-18927 */
    return fo_atlas_field_ex(parent, pw, ph, pmax, child, cw, ch, cmax, out, ow, oh, parent_zoom, fraction, 1.0,
                             0.0, 0.0, 0.0, palette_max_iter, 48, threads);
}

/* crop_colourise_impl, renderer.cpp:18930-19020 */
static int crop_colourise_impl(const float *src, int sw, int sh, uint8_t *out, int ow, int oh, double zoom,
                               int max_iter, double phase, double vocal, double instrumental, double pitch,
                               int threads, const int *interior, const uint8_t *accents) {
    if (!src || !out || !ok_dims(sw, sh) || !ok_dims(ow, oh) || !ok_iter(max_iter) || threads < 0 || threads > 4096
        || !isfinite(zoom) || zoom <= 0.0 || !ok_controls(phase, vocal, instrumental, pitch) || !ok_interior(interior))
        return 1;
    zoom = zoom > 1.0 ? zoom : 1.0;
    int size = 0;
    uint8_t *pal = palette_new(max_iter, phase, vocal, pitch, accents, &size);
    if (!pal) return 1;
    const float scale = (float)(size - 1) / (float)max_iter;
    axis_t ax = {0}, ay = {0};
    int ok = axis_new(&ax, sw, ow, zoom) && axis_new(&ay, sh, oh, zoom);
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            int inside = 0;
            const float s = sample_interior(src, sw, &ax, &ay, x, y, max_iter, &inside, 0.0);
            uint8_t *d = out + ((size_t)y * (size_t)ow + (size_t)x) * 3;
            if (inside) {
                for (int c = 0; c < 3; ++c) d[c] = interior ? (uint8_t)interior[c] : 0;
                continue;
            }
            const double sc = (double)s * (double)scale;
            const int k = sc >= (double)(size - 1) ? size - 1 : (!isfinite(sc) || sc <= 0.0 ? 0 : (int)sc);
            memcpy(d, pal + (size_t)k * 3, 3);
        }
    axis_free(&ax); axis_free(&ay); free(pal);
    return ok ? 0 : 1;
}

int fo_crop_colourise(const float *src, int sw, int sh, uint8_t *out, int ow, int oh, double zoom, int max_iter,
                      double phase, double vocal, double instrumental, double pitch, int threads) {
    return crop_colourise_impl(src, sw, sh, out, ow, oh, zoom, max_iter, phase, vocal, instrumental, pitch, threads,
                               NULL, NULL);
}
int fo_crop_colourise_interior(const float *src, int sw, int sh, uint8_t *out, int ow, int oh, double zoom,
                               int max_iter, double phase, double vocal, double instrumental, double pitch, int ir,
                               int ig, int ib, int threads) {
    const int interior[3] = {ir, ig, ib};
    return crop_colourise_impl(src, sw, sh, out, ow, oh, zoom, max_iter, phase, vocal, instrumental, pitch, threads,
                               interior, NULL);
}
int fo_crop_colourise_accents(const float *src, int sw, int sh, uint8_t *out, int ow, int oh, double zoom,
                              int max_iter, double phase, double vocal, double instrumental, double pitch,
                              const uint8_t *accents, int ir, int ig, int ib, int threads) {
    const int interior[3] = {ir, ig, ib};
    return crop_colourise_impl(src, sw, sh, out, ow, oh, zoom, max_iter, phase, vocal, instrumental, pitch, threads,
                               interior, accents);
}

/* atlas_colourise_impl, renderer.cpp:19137-19410 */
static int atlas_colourise_impl(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch,
                                int cmax, uint8_t *out, int ow, int oh, double parent_zoom, double fraction,
                                int palette_max_iter, double phase, double vocal, double instrumental, double pitch,
                                int threads, const int *interior, const uint8_t *accents) {
    if (!parent || !out || !ok_dims(pw, ph) || !ok_dims(ow, oh) || !ok_iter(pmax) || threads < 0 || threads > 4096
        || !isfinite(parent_zoom) || parent_zoom <= 0.0 || !isfinite(fraction) || fraction < 0.0 || fraction > 1.0
        || palette_max_iter < 0 || palette_max_iter > FO_MAX_ITER || !ok_controls(phase, vocal, instrumental, pitch)
        || !ok_interior(interior))
        return 1;
    const int use = child != NULL && fraction > 0.0;
    if (use && (!ok_dims(cw, ch) || !ok_iter(cmax))) return 1;
    const int eff = maxi(1, maxi(palette_max_iter, maxi(pmax, use ? cmax : 0)));
    int size = 0;
    uint8_t *pal = palette_new(eff, phase, vocal, pitch, accents, &size);
    if (!pal) return 1;
    const float scale = (float)(size - 1) / (float)eff;
    const int vw = use ? maxi(1, (int)lround((double)ow * fraction)) : 0;
    const int vh = use ? maxi(1, (int)lround((double)oh * fraction)) : 0;
    const int cl = (ow - vw) / 2, ct = (oh - vh) / 2;
    const int fe = use ? mini(48, mini(vw / 8, vh / 8)) : 0;
    const int full = use && fraction >= 0.999999;
    axis_t px = {0}, py = {0}, cx = {0}, cy = {0};
    int ok = 1;
    if (!full) ok = axis_new(&px, pw, ow, parent_zoom) && axis_new(&py, ph, oh, parent_zoom);
    if (ok && use) ok = axis_new(&cx, cw, full ? ow : vw, 1.0) && axis_new(&cy, ch, full ? oh : vh, 1.0);
    float *ex = NULL, *ey = NULL;
    if (ok && use && !full && fe >= 2) { ex = seam(vw, fe); ey = seam(vh, fe); ok = ex && ey; }
    const float inner = (float)eff;
    for (int y = 0; ok && y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            uint8_t *d = out + ((size_t)y * (size_t)ow + (size_t)x) * 3;
            int pin = 0, cin = 0;
            if (full) {
                const float s = sample_interior(child, cw, &cx, &cy, x, y, cmax, &cin, 0.0);
                colour(cin ? inner : s, eff, pal, size, scale, interior, d);
                continue;
            }
            if (!use || y < ct || y >= ct + vh || x < cl || x >= cl + vw) {
                const float s = sample_interior(parent, pw, &px, &py, x, y, pmax, &pin, 0.0);
                colour(pin ? inner : s, eff, pal, size, scale, interior, d);
                continue;
            }
            const int kx = x - cl, ky = y - ct;
            const float cs = sample_interior(child, cw, &cx, &cy, kx, ky, cmax, &cin, 0.0);
            const float a = fe >= 2 ? (ex[kx] < ey[ky] ? ex[kx] : ey[ky]) : 1.0F;
            if (a >= 0.999999F) {
                colour(cin ? inner : cs, eff, pal, size, scale, interior, d);
                continue;
            }
            const float ps = sample_interior(parent, pw, &px, &py, x, y, pmax, &pin, 0.0);
            uint8_t prgb[3], crgb[3];
            colour(pin ? inner : ps, eff, pal, size, scale, interior, prgb);
            colour(cin ? inner : cs, eff, pal, size, scale, interior, crgb);
            for (int c = 0; c < 3; ++c) d[c] = byte_round((double)prgb[c] * (1.0 - a) + (double)crgb[c] * a);
        }
    free(ex); free(ey); free(pal);
    axis_free(&px); axis_free(&py); axis_free(&cx); axis_free(&cy);
    return ok ? 0 : 1;
}

int fo_atlas_colourise(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch, int cmax,
                       uint8_t *out, int ow, int oh, double parent_zoom, double fraction, int palette_max_iter,
                       double phase, double vocal, double instrumental, double pitch, int threads) {
    return atlas_colourise_impl(parent, pw, ph, pmax, child, cw, ch, cmax, out, ow, oh, parent_zoom, fraction,
                                palette_max_iter, phase, vocal, instrumental, pitch, threads, NULL, NULL);
}
int fo_atlas_colourise_interior(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch,
                                int cmax, uint8_t *out, int ow, int oh, double parent_zoom, double fraction,
                                int palette_max_iter, double phase, double vocal, double instrumental, double pitch,
                                int ir, int ig, int ib, int threads) {
    const int interior[3] = {ir, ig, ib};
    return atlas_colourise_impl(parent, pw, ph, pmax, child, cw, ch, cmax, out, ow, oh, parent_zoom, fraction,
                                palette_max_iter, phase, vocal, instrumental, pitch, threads, interior, NULL);
}
int fo_atlas_colourise_accents(const float *parent, int pw, int ph, int pmax, const float *child, int cw, int ch,
                               int cmax, uint8_t *out, int ow, int oh, double parent_zoom, double fraction,
                               int palette_max_iter, double phase, double vocal, double instrumental, double pitch,
                               const uint8_t *accents, int ir, int ig, int ib, int threads) {
    if (!accents) return 1;
    const int interior[3] = {ir, ig, ib};
    return atlas_colourise_impl(parent, pw, ph, pmax, child, cw, ch, cmax, out, ow, oh, parent_zoom, fraction,
                                palette_max_iter, phase, vocal, instrumental, pitch, threads, interior, accents);
}
