/*
 * fav_oracle.c -- plain-C restatement of the reference's scalar FP64 deep
 * Mandelbrot path.  TEST INFRASTRUCTURE ONLY (see fav_oracle.h).
 *
 * Every function names the renderer.cpp lines whose behaviour it restates.
 * The order of floating-point operations follows the reference exactly, since
 * escape counts of chaotic pixels depend on every rounding.
 */
#define _POSIX_C_SOURCE 200809L
#include "fav_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <gmp.h>
#include <mpfr.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants (renderer.cpp:66-97, 4811) ------------------------------ */
#define MAX_SAFE_BLA_LENGTH 64
#define MAX_SAFE_LINEAR_BLA_LENGTH 4096
#define MAX_SAFE_DEEP_LINEAR_BLA_LENGTH 1024
static const long double LOG_TWO_L = 0.693147180559945309417232121458176568L;
static const long double LOG_TEN_L = 2.302585092994045684017991454684364208L;
static const double TWO_PI_D = 6.283185307179586476925286766559005768;

static void fo_set_error(char *error, int capacity, const char *fmt, ...) {
    if (!error || capacity <= 0) return;
    va_list args;
    va_start(args, fmt);
    vsnprintf(error, (size_t)capacity, fmt, args);
    va_end(args);
}

/* ---- FloatExp (renderer.cpp:7532-7592) --------------------------------- */
typedef struct { double m; int e; } fe_t;
typedef struct { fe_t re, im; } fec_t;
typedef struct { double re, im; int e; } sc_t;
typedef struct { double m; int e; } sn_t; /* ScaledNorm */

static inline int sat_exp(long long v) { /* renderer.cpp:7532-7540 */
    if (v > INT_MAX) return INT_MAX;
    if (v < INT_MIN) return INT_MIN;
    return (int)v;
}

static inline fe_t fe_make(double m, int e) { fe_t r = {m, e}; return r; }
static const fe_t FE_ZERO = {0.0, 0};

/* FloatExp::from_parts, renderer.cpp:7551-7560. */
static inline fe_t fe_from_parts(double value, int exponent) {
    if (value == 0.0 || isnan(value)) return fe_make(value, 0);
    int shift = 0;
    const double normalized = frexp(value, &shift);
    return fe_make(normalized, sat_exp((long long)exponent + (long long)shift));
}

/* FloatExp::from_mpfr, renderer.cpp:7570-7579. */
static fe_t fe_from_mpfr(mpfr_srcptr value) {
    long exponent = 0;
    const double mantissa = mpfr_get_d_2exp(&exponent, value, MPFR_RNDN);
    if (!isfinite(mantissa)) return fe_make(mantissa, 0);
    if (exponent > INT_MAX) return fe_make(copysign(INFINITY, mantissa), 0);
    if (exponent < INT_MIN) return FE_ZERO;
    return fe_from_parts(mantissa, (int)exponent);
}

static inline int fe_finite(fe_t v) { return isfinite(v.m); }
static inline int fe_zero(fe_t v) { return v.m == 0.0; }

/* escape_radius_squared_float_exp, renderer.cpp:7594-7598. */
static inline fe_t fe_escape_r2(int mode) {
    /* mode 2: radius 1e6 (R^2 = 1e12), a mode of neither the reference nor
       the HIP library, kept to show what the reference's norm does there */
    return mode == 1 ? fe_from_parts(1.0e8, 0) : (mode == 2 ? fe_from_parts(1.0e12, 0) : fe_from_parts(4.0, 0));
}

/* ---- FloatExp arithmetic (renderer.cpp:8190-8349) ---------------------- */
static inline fe_t fe_neg(fe_t v) { return fe_make(-v.m, v.e); }

static fe_t fe_add(fe_t a, fe_t b) { /* renderer.cpp:8198-8235 */
    if (fe_zero(a)) return b;
    if (fe_zero(b)) return a;
    const fe_t *larger = &a, *smaller = &b;
    if (b.e > a.e) { larger = &b; smaller = &a; }
    const long long diff = (long long)larger->e - (long long)smaller->e;
    if (diff > 60) return *larger;
    const double mantissa = larger->m + ldexp(smaller->m, -(int)diff);
    if (mantissa == 0.0 || !isfinite(mantissa)) return fe_from_parts(mantissa, larger->e);
    const double magnitude = fabs(mantissa);
    if (magnitude >= 1.0) return fe_make(mantissa * 0.5, sat_exp((long long)larger->e + 1));
    if (magnitude >= 0.5) return fe_make(mantissa, larger->e);
    int shift = 0;
    const double n = frexp(mantissa, &shift);
    return fe_make(n, sat_exp((long long)larger->e + (long long)shift));
}

static inline fe_t fe_sub(fe_t a, fe_t b) { return fe_add(a, fe_neg(b)); }

static fe_t fe_mul(fe_t a, fe_t b) { /* renderer.cpp:8241-8266 */
    if (fe_zero(a) || fe_zero(b)) return FE_ZERO;
    const double mantissa = a.m * b.m;
    if (!isfinite(mantissa)) return fe_from_parts(mantissa, 0);
    const double magnitude = fabs(mantissa);
    const long long sum = (long long)a.e + (long long)b.e;
    if (magnitude >= 1.0) return fe_make(mantissa * 0.5, sat_exp(sum + 1));
    if (magnitude < 0.5) return fe_make(mantissa * 2.0, sat_exp(sum - 1));
    return fe_make(mantissa, sat_exp(sum));
}

/* fe_mul(FloatExp, double), renderer.cpp:8268-8270. */
static inline fe_t fe_muld(fe_t a, double b) { return fe_mul(a, fe_from_parts(b, 0)); }

static fe_t fe_div(fe_t a, fe_t b) { /* renderer.cpp:8272-8293 */
    if (fe_zero(a)) return FE_ZERO;
    const double mantissa = a.m / b.m;
    if (!isfinite(mantissa)) return fe_from_parts(mantissa, 0);
    const double magnitude = fabs(mantissa);
    const long long diff = (long long)a.e - (long long)b.e;
    if (magnitude >= 1.0) return fe_make(mantissa * 0.5, sat_exp(diff + 1));
    if (magnitude < 0.5) return fe_make(mantissa * 2.0, sat_exp(diff - 1));
    return fe_make(mantissa, sat_exp(diff));
}

static inline fe_t fe_sqr(fe_t v) { return fe_mul(v, v); }

static fe_t fe_sqrt(fe_t v) { /* renderer.cpp:8299-8316 */
    if (fe_zero(v)) return v;
    long long exponent = v.e;
    double mantissa = v.m;
    if (mantissa < 0.0 || !isfinite(mantissa)) return fe_from_parts(sqrt(mantissa), 0);
    if (exponent % 2 != 0) {
        mantissa *= 2.0;
        --exponent;
        return fe_make(sqrt(mantissa) * 0.5, sat_exp(exponent / 2 + 1));
    }
    return fe_make(sqrt(mantissa), sat_exp(exponent / 2));
}

static int fe_compare(fe_t a, fe_t b) { /* renderer.cpp:8318-8339 */
    if (!fe_finite(a)) return fe_finite(b) ? 1 : 0;
    if (!fe_finite(b)) return -1;
    if (a.m == 0.0) return b.m == 0.0 ? 0 : -1;
    if (b.m == 0.0) return a.m > 0.0 ? 1 : -1;
    if (a.m < 0.0 && b.m >= 0.0) return -1;
    if (a.m >= 0.0 && b.m < 0.0) return 1;
    const int negative = a.m < 0.0;
    if (a.e != b.e) {
        const int result = a.e < b.e ? -1 : 1;
        return negative ? -result : result;
    }
    return (a.m > b.m) - (a.m < b.m);
}

static inline fe_t fe_min(fe_t a, fe_t b) { return fe_compare(a, b) <= 0 ? a : b; } /* 11336 */

static inline fe_t fe_norm_squared(fe_t re, fe_t im) { return fe_add(fe_sqr(re), fe_sqr(im)); }

static long double fe_log(fe_t v) { /* renderer.cpp:8345-8349 */
    if (v.m == 0.0) return -INFINITY;
    return logl(fabsl((long double)v.m)) + (long double)v.e * LOG_TWO_L;
}

/* FloatExpComplex ops, renderer.cpp:8351-8381. */
static inline fec_t fec_make(fe_t re, fe_t im) { fec_t r = {re, im}; return r; }
static inline fec_t fec_add(fec_t a, fec_t b) { return fec_make(fe_add(a.re, b.re), fe_add(a.im, b.im)); }
static inline fec_t fec_neg(fec_t v) { return fec_make(fe_neg(v.re), fe_neg(v.im)); }
static inline fec_t fec_sub(fec_t a, fec_t b) { return fec_add(a, fec_neg(b)); }
static inline fec_t fec_mul(fec_t a, fec_t b) {
    return fec_make(fe_sub(fe_mul(a.re, b.re), fe_mul(a.im, b.im)),
                    fe_add(fe_mul(a.re, b.im), fe_mul(a.im, b.re)));
}
static inline fec_t fec_mulf(fec_t a, fe_t b) { return fec_make(fe_mul(a.re, b), fe_mul(a.im, b)); }
static inline fe_t fec_norm_squared(fec_t v) { return fe_norm_squared(v.re, v.im); }
static inline int fec_finite(fec_t v) { return fe_finite(v.re) && fe_finite(v.im); }

/* fec_escape_margin_with_delta, renderer.cpp:8383-8399. */
static fe_t fec_escape_margin_with_delta(fec_t reference, fec_t delta, int mode) {
    const fe_t cross = fe_muld(fe_add(fe_mul(reference.re, delta.re),
                                      fe_mul(reference.im, delta.im)), 2.0);
    const fe_t reference_margin = fe_sub(fec_norm_squared(reference), fe_escape_r2(mode));
    return fe_add(fe_add(reference_margin, cross), fec_norm_squared(delta));
}

/* ---- ScaledComplex (renderer.cpp:7605-7736) ---------------------------- */
static inline sc_t sc_make(double re, double im, int e) { sc_t r = {re, im, e}; return r; }
static const sc_t SC_ZERO = {0.0, 0.0, 0};

static void sc_normalize(sc_t *v) { /* renderer.cpp:7610-7655 */
    const double ar = fabs(v->re), ai = fabs(v->im);
    const double magnitude = (ar < ai) ? ai : ar; /* std::max(ar, ai) */
    if (magnitude == 0.0) { v->re = 0.0; v->im = 0.0; v->e = 0; return; }
    if (magnitude >= 2.0) {
        int shift = 0;
        (void)frexp(magnitude, &shift);
        v->re = ldexp(v->re, -shift);
        v->im = ldexp(v->im, -shift);
        v->e = sat_exp((long long)v->e + (long long)shift);
        return;
    }
    if (magnitude >= 1.0) {
        v->re *= 0.5; v->im *= 0.5;
        v->e = sat_exp((long long)v->e + 1);
        return;
    }
    if (magnitude >= 0.5) return;
    if (magnitude >= 0.25) {
        v->re *= 2.0; v->im *= 2.0;
        v->e = sat_exp((long long)v->e - 1);
        return;
    }
    int shift = 0;
    (void)frexp(magnitude, &shift);
    v->re = ldexp(v->re, -shift);
    v->im = ldexp(v->im, -shift);
    v->e = sat_exp((long long)v->e + (long long)shift);
}

/* ScaledComplex::from_float_exp, renderer.cpp:7657-7674. */
static sc_t sc_from_fe(fe_t re, fe_t im) {
    if (fe_zero(re) && fe_zero(im)) return SC_ZERO;
    const int common = re.e > im.e ? re.e : im.e;
    const long long rd = (long long)re.e - (long long)common;
    const long long id = (long long)im.e - (long long)common;
    sc_t r = sc_make(rd < -1074 ? 0.0 : ldexp(re.m, (int)rd),
                     id < -1074 ? 0.0 : ldexp(im.m, (int)id), common);
    sc_normalize(&r);
    return r;
}

static sc_t sc_add(sc_t a, sc_t b) { /* renderer.cpp:7676-7697 */
    if (a.re == 0.0 && a.im == 0.0) return b;
    if (b.re == 0.0 && b.im == 0.0) return a;
    const sc_t *larger = &a, *smaller = &b;
    if (b.e > a.e) { larger = &b; smaller = &a; }
    const long long diff = (long long)larger->e - (long long)smaller->e;
    if (diff > 60) return *larger;
    const int d = (int)diff;
    sc_t r = sc_make(larger->re + ldexp(smaller->re, -d), larger->im + ldexp(smaller->im, -d), larger->e);
    sc_normalize(&r);
    return r;
}

static inline sc_t sc_neg(sc_t v) { return sc_make(-v.re, -v.im, v.e); }
static inline sc_t sc_conj(sc_t v) { return sc_make(v.re, -v.im, v.e); }
static inline sc_t sc_sub(sc_t a, sc_t b) { return sc_add(a, sc_neg(b)); }

static sc_t sc_mul(sc_t a, sc_t b) { /* renderer.cpp:7711-7722 */
    if ((a.re == 0.0 && a.im == 0.0) || (b.re == 0.0 && b.im == 0.0)) return SC_ZERO;
    sc_t r = sc_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re,
                     sat_exp((long long)a.e + (long long)b.e));
    sc_normalize(&r);
    return r;
}

static inline sc_t sc_double(sc_t v) { /* renderer.cpp:7724-7733 */
    if (v.re == 0.0 && v.im == 0.0) return SC_ZERO;
    return sc_make(v.re, v.im, sat_exp((long long)v.e + 1));
}

static inline int sc_finite(sc_t v) { return isfinite(v.re) && isfinite(v.im); }

static sn_t sc_norm_squared(sc_t v) { /* renderer.cpp:7743-7774 */
    const sn_t inf = {INFINITY, INT_MAX};
    if (!sc_finite(v)) return inf;
    const double norm = v.re * v.re + v.im * v.im;
    if (norm == 0.0) { sn_t z = {0.0, 0}; return z; }
    if (!isfinite(norm)) return inf;
    const long long se = 2LL * (long long)v.e;
    if (norm >= 1.0) {
        const long long e = se + 1;
        if (e > INT_MAX) return inf;
        sn_t r = {norm * 0.5, (int)e};
        return r;
    }
    if (norm < 0.5) {
        const long long e = se - 1;
        if (e < INT_MIN) { sn_t z = {0.0, 0}; return z; }
        sn_t r = {norm * 2.0, (int)e};
        return r;
    }
    if (se > INT_MAX) return inf;
    if (se < INT_MIN) { sn_t z = {0.0, 0}; return z; }
    sn_t r = {norm, (int)se};
    return r;
}

static int sc_compare_norm(sn_t a, sn_t b) { /* renderer.cpp:7776-7784 */
    if (!isfinite(a.m)) return isfinite(b.m) ? 1 : 0;
    if (!isfinite(b.m)) return -1;
    if (a.m == 0.0 && b.m == 0.0) return 0;
    if (a.m == 0.0) return -1;
    if (b.m == 0.0) return a.m > 0.0 ? 1 : -1;
    if (a.e != b.e) return a.e < b.e ? -1 : 1;
    return (a.m > b.m) - (a.m < b.m);
}

static inline int sc_outside_escape(sn_t norm, int mode) { /* renderer.cpp:7786-7793 */
    const fe_t t = fe_escape_r2(mode);
    sn_t th = {t.m, t.e};
    return sc_compare_norm(norm, th) > 0;
}

static inline double sc_re_double(sc_t v) { return ldexp(v.re, v.e); }
static inline double sc_im_double(sc_t v) { return ldexp(v.im, v.e); }

static double scaled_norm_to_double(sn_t n) { /* renderer.cpp:7803-7813 */
    if (n.m == 0.0) return 0.0;
    if (!isfinite(n.m)) return DBL_MAX;
    const double value = ldexp(n.m, n.e);
    return isfinite(value) ? (0.0 < value ? value : 0.0) : DBL_MAX;
}

static sn_t scaled_norm_from_double(double value) { /* renderer.cpp:7815-7824 */
    if (!(value > 0.0) || !isfinite(value)) {
        sn_t r = {value > 0.0 ? DBL_MAX : 0.0, 0};
        return r;
    }
    int e = 0;
    const double m = frexp(value, &e);
    sn_t r = {m, e};
    return r;
}

/* sc_component_as_float_exp + sc_escape_margin_with_delta,
 * renderer.cpp:8401-8431. */
static inline fe_t sc_part(sc_t v, int imaginary) {
    return fe_from_parts(imaginary ? v.im : v.re, v.e);
}
static fe_t sc_escape_margin_with_delta(sc_t reference, sc_t delta, int mode) {
    return fec_escape_margin_with_delta(fec_make(sc_part(reference, 0), sc_part(reference, 1)),
                                        fec_make(sc_part(delta, 0), sc_part(delta, 1)), mode);
}

/* sc_norm_squared_with_delta, renderer.cpp:8451-8461. */
static inline sn_t sc_norm_with_margin(fe_t margin, int mode) {
    const fe_t n = fe_add(fe_escape_r2(mode), margin);
    sn_t r = {n.m, n.e};
    return r;
}

/* ---- render planes (renderer.cpp:7826-7896, 8075-8130) ----------------- */
static int render_plane_de(sc_t total, sc_t derivative, fe_t pixel_spacing,
                           double *out_re, double *out_im) {
    const double cm = hypot(total.re, total.im);
    if (!(cm > 0.0) || !isfinite(cm)) return 0;
    /* LOG_TWO is long double in the reference, so the sum is formed in long
     * double and rounded once to double. */
    const double lm = (double)((long double)log(cm) + (long double)(double)total.e * LOG_TWO_L);
    if (!(lm > 0.0) || !isfinite(lm)) return 0;
    const double ur = total.re / cm;
    const double ui = total.im / cm;
    const sc_t unit = sc_from_fe(fe_from_parts(ur, 0), fe_from_parts(ui, 0));
    const sc_t spacing = sc_from_fe(pixel_spacing, FE_ZERO);
    const sc_t den = sc_mul(unit, sc_mul(sc_conj(derivative), spacing));
    const double den2 = den.re * den.re + den.im * den.im;
    if (!(den2 > 0.0) || !isfinite(den2)) return 0;
    sc_t r = sc_make(cm * lm * den.re / den2, -cm * lm * den.im / den2,
                     sat_exp((long long)total.e - (long long)den.e));
    sc_normalize(&r);
    *out_re = sc_re_double(r);
    *out_im = sc_im_double(r);
    return isfinite(*out_re) && isfinite(*out_im);
}

static double render_plane_phase(sc_t v) { /* renderer.cpp:8075-8082 */
    const double re = sc_re_double(v), im = sc_im_double(v);
    if (!isfinite(re) || !isfinite(im)) return 0.0;
    double phase = atan2(im, re) / TWO_PI_D;
    phase -= floor(phase);
    return isfinite(phase) ? phase : 0.0;
}

static void clear_planes_pixel(FractalRenderPlanes *p, size_t i, int max_iter) { /* 8084-8096 */
    if (!p) return;
    p->orbit_iteration[i] = max_iter;
    p->phase[i] = 0.0; p->de_x[i] = 0.0; p->de_y[i] = 0.0;
    p->test1[i] = 0.0; p->test2[i] = 0.0;
}

static inline int kalles_raw_iteration(int iteration) { return iteration - 1 > 0 ? iteration - 1 : 0; }

static void store_planes_escape(FractalRenderPlanes *p, size_t i, int iteration, sc_t total,
                                sn_t test1, sn_t test2, sc_t derivative, fe_t spacing) {
    /* renderer.cpp:8106-8135 */
    if (!p) return;
    p->orbit_iteration[i] = iteration;
    p->phase[i] = render_plane_phase(total);
    p->de_x[i] = 0.0; p->de_y[i] = 0.0;
    double dx = 0.0, dy = 0.0;
    if (render_plane_de(total, derivative, spacing, &dx, &dy)) { p->de_x[i] = dx; p->de_y[i] = dy; }
    p->test1[i] = scaled_norm_to_double(test1);
    p->test2[i] = scaled_norm_to_double(test2);
}

/* smooth_escape_value / smooth_escape_scaled / encode_render_value,
 * renderer.cpp:8170-8188, 119-125. */
static float smooth_escape_scaled(int iteration, sn_t norm, double bias) {
    long double value = (long double)iteration;
    if (norm.m != 0.0) {
        const long double lm = 0.5L * (logl((long double)norm.m) + (long double)norm.e * LOG_TWO_L);
        if (lm > 0.0L && isfinite(lm)) value = (long double)iteration - logl(lm) / LOG_TWO_L;
    }
    return (float)(value - (long double)bias);
}
static inline float encode_iteration(int iteration, double bias) {
    return (float)((long double)iteration - (long double)bias);
}

/* ---- reference context ------------------------------------------------- */
typedef struct { sc_t c[9]; fe_t r2; int length; } cubic_step_t;       /* FastBlaStep */
typedef struct { sc_t A, B; fe_t r2; int length; } linear_step_t;      /* LinearBlaStep */
typedef struct { fec_t A, B, C, D, E, F, G, H, I; fe_t r2; int length; } bla_builder_t; /* BlaStep */
typedef struct { fec_t A, B; fe_t r2; int length; } linear_builder_t;

typedef struct { int count; size_t *offset; size_t *size; void *data; } levels_t;

struct fo_reference {
    /* compact orbit, renderer.cpp:8881-8890 */
    int orbit_size;
    sc_t *scaled;
    double *real_double, *imag_double;
    fec_t *fast_orbit; /* builder orbit, kept for clones */
    int fast_orbit_size;
    /* BlaLevels, renderer.cpp:8797-8879 */
    int cubic_levels;
    size_t *cubic_offset, *cubic_size;
    cubic_step_t *cubic;
    int linear_levels[2];
    size_t *linear_offset[2], *linear_size[2];
    linear_step_t *linear[2];
    fe_t input_radius, deep_input_radius;
    int map_end;
    /* ImageSeries, renderer.cpp:8591-8597 */
    int series_enabled, series_order, series_iteration;
    fe_t series_radius_squared;
    sc_t *series_coefficients;
    int series_count;
    int requested_max_iter, requested_series_order;
    int escape_radius_mode, coordinate_mode;
    int bla_epsilon_exponent;    /* BLA validity radius eps = 2^this * |Z| (the reference's -38) */
    long precision_bits;
    /* alternate formulas (renderer.cpp:8507-8565, 8881-8905) */
    int formula;                 /* 0 Mandelbrot, 1 Julia, 2 burning ship, 3 tricorn */
    sc_t parameter;              /* c (Julia constant or the centre) */
    fe_t *margins;               /* MPFR escape margin |Z|^2 - R^2 per orbit entry */
    int alt_levels;
    size_t *alt_offset, *alt_size;
    struct alt_step *alt;
    fe_t alt_input_radius_squared;
    int alt_start;
};

void fo_reference_destroy(fo_reference *r) {
    if (!r) return;
    free(r->scaled); free(r->real_double); free(r->imag_double); free(r->fast_orbit);
    free(r->cubic_offset); free(r->cubic_size); free(r->cubic);
    for (int t = 0; t < 2; ++t) { free(r->linear_offset[t]); free(r->linear_size[t]); free(r->linear[t]); }
    free(r->series_coefficients);
    free(r->margins); free(r->alt_offset); free(r->alt_size); free(r->alt);
    free(r);
}

/* parse_zoom_float_exp, renderer.cpp:11892-11915. */
static int parse_zoom_fe(const char *text, long precision, fe_t *out) {
    mpfr_t value;
    mpfr_init2(value, precision);
    const int status = mpfr_set_str(value, text, 10, MPFR_RNDN);
    if (status != 0 || mpfr_sgn(value) <= 0 || !mpfr_number_p(value)) { mpfr_clear(value); return 0; }
    const fe_t result = fe_from_mpfr(value);
    mpfr_clear(value);
    if (!fe_finite(result) || fe_zero(result)) return 0;
    const long double l10 = fe_log(result) / LOG_TEN_L;
    if (!isfinite(l10) || l10 < -300.0L || l10 > 9800.0L) return 0;
    *out = result;
    return 1;
}

static inline double viewport_height_factor(int mode) { return mode == 1 ? 4.0 : 2.8; } /* 153-158 */

/* escape_radius_squared_double, renderer.cpp:160-168 (R = 2 or 10000). */
static inline double mode_r2_double(int mode) { return mode == 1 ? 1.0e8 : (mode == 2 ? 1.0e12 : 4.0); }

static inline double pixel_axis_offset(int index, int dimension, int mode) { /* 142-151 */
    const double origin = mode == 1 ? (double)(dimension / 2) : (double)(dimension - 1) / 2.0;
    return (double)index - origin;
}

static void stabilize_alternate_cycle(fo_reference *r);

/* make_reference_orbit, renderer.cpp:10640-10922: MPFR iteration of the
 * formula at the handle's precision; Julia starts from the repelling fixed
 * point when the centre is within 10^-digits of it (10736-10815). */
static int make_orbit(fo_reference *r, const char *x, const char *y, const char *zoom_text,
                      int max_iter, int precision_bits, const char *julia_re, const char *julia_im,
                      char *error, int cap) {
    precision_bits = precision_bits > 128 ? precision_bits : 128;
    r->precision_bits = precision_bits;
    const int formula = r->formula;
    mpfr_t cx, cy, vz, vr, zr, zi, nr, ni, tmp, norm, margin, pr, pi, ar, ai;
    mpfr_t dre, dim, dmag, rre, rim, are, aim, bre, bim, anorm, bnorm, dist, mult, tol, vlog;
    mpfr_inits2(precision_bits, cx, cy, vz, vr, zr, zi, nr, ni, tmp, norm, margin, pr, pi, ar, ai,
                dre, dim, dmag, rre, rim, are, aim, bre, bim, anorm, bnorm, dist, mult, tol, vlog, (mpfr_ptr)0);
    int ok = 0;
    if (mpfr_set_str(cx, x, 10, MPFR_RNDN) != 0 || mpfr_set_str(cy, y, 10, MPFR_RNDN) != 0) {
        fo_set_error(error, cap, "invalid MPFR Mandelbrot centre");
        goto done;
    }
    const int have_zoom = zoom_text && mpfr_set_str(vz, zoom_text, 10, MPFR_RNDN) == 0 && mpfr_sgn(vz) > 0;
    if (have_zoom) {
        mpfr_set_d(vr, viewport_height_factor(r->coordinate_mode), MPFR_RNDN);
        mpfr_div(vr, vr, vz, MPFR_RNDN);
        r->input_radius = fe_from_mpfr(vr);
    } else {
        r->input_radius = fe_from_parts(1.0, 0);
    }
    if (formula == 1) {
        if (mpfr_set_str(pr, julia_re, 10, MPFR_RNDN) != 0 || mpfr_set_str(pi, julia_im, 10, MPFR_RNDN) != 0) {
            fo_set_error(error, cap, "invalid MPFR Julia constant");
            goto done;
        }
        mpfr_set(zr, cx, MPFR_RNDN);
        mpfr_set(zi, cy, MPFR_RNDN);
    } else {
        mpfr_set(pr, cx, MPFR_RNDN);
        mpfr_set(pi, cy, MPFR_RNDN);
        mpfr_set_zero(zr, 0);
        mpfr_set_zero(zi, 0);
    }
    r->parameter = sc_from_fe(fe_from_mpfr(pr), fe_from_mpfr(pi));
    int fixed_point = 0;
    if (formula == 1) {
        /* roots of z^2 - z + c: (1 +- sqrt(1 - 4c))/2, the larger |z| repels */
        mpfr_set_ui(tmp, 4, MPFR_RNDN);
        mpfr_mul(dre, pr, tmp, MPFR_RNDN);
        mpfr_ui_sub(dre, 1, dre, MPFR_RNDN);
        mpfr_mul(dim, pi, tmp, MPFR_RNDN);
        mpfr_neg(dim, dim, MPFR_RNDN);
        mpfr_mul(norm, dre, dre, MPFR_RNDN);
        mpfr_mul(tmp, dim, dim, MPFR_RNDN);
        mpfr_add(norm, norm, tmp, MPFR_RNDN);
        mpfr_sqrt(dmag, norm, MPFR_RNDN);
        mpfr_add(rre, dmag, dre, MPFR_RNDN);
        mpfr_div_2ui(rre, rre, 1, MPFR_RNDN);
        mpfr_sqrt(rre, rre, MPFR_RNDN);
        mpfr_sub(rim, dmag, dre, MPFR_RNDN);
        mpfr_div_2ui(rim, rim, 1, MPFR_RNDN);
        mpfr_sqrt(rim, rim, MPFR_RNDN);
        if (mpfr_sgn(dim) < 0) mpfr_neg(rim, rim, MPFR_RNDN);
        mpfr_add_ui(are, rre, 1, MPFR_RNDN);
        mpfr_div_2ui(are, are, 1, MPFR_RNDN);
        mpfr_div_2ui(aim, rim, 1, MPFR_RNDN);
        mpfr_set_ui(tmp, 1, MPFR_RNDN);
        mpfr_sub(bre, tmp, rre, MPFR_RNDN);
        mpfr_div_2ui(bre, bre, 1, MPFR_RNDN);
        mpfr_neg(bim, aim, MPFR_RNDN);
        mpfr_mul(anorm, are, are, MPFR_RNDN);
        mpfr_mul(tmp, aim, aim, MPFR_RNDN);
        mpfr_add(anorm, anorm, tmp, MPFR_RNDN);
        mpfr_mul(bnorm, bre, bre, MPFR_RNDN);
        mpfr_mul(tmp, bim, bim, MPFR_RNDN);
        mpfr_add(bnorm, bnorm, tmp, MPFR_RNDN);
        if (mpfr_cmp(anorm, bnorm) >= 0) {
            mpfr_set(nr, are, MPFR_RNDN); mpfr_set(ni, aim, MPFR_RNDN); mpfr_set(norm, anorm, MPFR_RNDN);
        } else {
            mpfr_set(nr, bre, MPFR_RNDN); mpfr_set(ni, bim, MPFR_RNDN); mpfr_set(norm, bnorm, MPFR_RNDN);
        }
        mpfr_mul_ui(mult, norm, 4, MPFR_RNDN);
        /* the zoom was validated, so viewport_zoom holds it here */
        mpfr_log10(vlog, vz, MPFR_RNDN);
        const double log_zoom = mpfr_get_d(vlog, MPFR_RNDN);
        int digits = (int)ceil(log_zoom > 0.0 ? log_zoom : 0.0) + 16;
        if (digits < 32) digits = 32;
        mpfr_set_ui(tol, 10, MPFR_RNDN);
        mpfr_pow_si(tol, tol, -(long)digits, MPFR_RNDN);
        mpfr_sub(tmp, cx, nr, MPFR_RNDN);
        mpfr_mul(dist, tmp, tmp, MPFR_RNDN);
        mpfr_sub(tmp, cy, ni, MPFR_RNDN);
        mpfr_mul(norm, tmp, tmp, MPFR_RNDN);
        mpfr_add(dist, dist, norm, MPFR_RNDN);
        mpfr_mul(norm, tol, tol, MPFR_RNDN);
        if (mpfr_cmp_ui(mult, 1) > 0 && mpfr_cmp(dist, norm) <= 0) {
            mpfr_set(zr, nr, MPFR_RNDN);
            mpfr_set(zi, ni, MPFR_RNDN);
            fixed_point = 1;
        }
    }
    r->fast_orbit = (fec_t *)malloc(sizeof(fec_t) * ((size_t)max_iter + 1));
    r->margins = (fe_t *)malloc(sizeof(fe_t) * ((size_t)max_iter + 1));
    if (!r->fast_orbit || !r->margins) { fo_set_error(error, cap, "out of memory"); goto done; }
    fe_t ore = fe_from_mpfr(zr), oim = fe_from_mpfr(zi);
    int finite_size = 0;
    const double r2 = mode_r2_double(r->escape_radius_mode);
    for (int i = 0; i <= max_iter; ++i) {
        if (!fe_finite(ore) || !fe_finite(oim)) break;
        r->fast_orbit[finite_size] = fec_make(ore, oim);
        mpfr_mul(norm, zr, zr, MPFR_RNDN);
        mpfr_mul(tmp, zi, zi, MPFR_RNDN);
        mpfr_add(norm, norm, tmp, MPFR_RNDN);
        mpfr_set_d(tmp, r2, MPFR_RNDN);
        mpfr_sub(margin, norm, tmp, MPFR_RNDN);
        r->margins[finite_size] = fe_from_mpfr(margin);
        ++finite_size;
        if (fixed_point) continue;
        if (formula == 2) {
            mpfr_abs(ar, zr, MPFR_RNDN);
            mpfr_abs(ai, zi, MPFR_RNDN);
            mpfr_mul(nr, ar, ar, MPFR_RNDN);
            mpfr_mul(tmp, ai, ai, MPFR_RNDN);
            mpfr_sub(nr, nr, tmp, MPFR_RNDN);
            mpfr_add(nr, nr, pr, MPFR_RNDN);
            mpfr_mul(ni, ar, ai, MPFR_RNDN);
            mpfr_mul_ui(ni, ni, 2, MPFR_RNDN);
            mpfr_add(ni, ni, pi, MPFR_RNDN);
        } else {
            mpfr_mul(nr, zr, zr, MPFR_RNDN);
            mpfr_mul(tmp, zi, zi, MPFR_RNDN);
            mpfr_sub(nr, nr, tmp, MPFR_RNDN);
            mpfr_mul(ni, zr, zi, MPFR_RNDN);
            mpfr_mul_ui(ni, ni, 2, MPFR_RNDN);
            if (formula == 3) mpfr_neg(ni, ni, MPFR_RNDN);
            mpfr_add(nr, nr, pr, MPFR_RNDN);
            mpfr_add(ni, ni, pi, MPFR_RNDN);
        }
        mpfr_set(zr, nr, MPFR_RNDN);
        mpfr_set(zi, ni, MPFR_RNDN);
        ore = fe_from_mpfr(zr);
        oim = fe_from_mpfr(zi);
    }
    if (finite_size == 0) { fo_set_error(error, cap, "reference orbit lost finite state at iteration zero"); goto done; }
    r->fast_orbit_size = finite_size;
    r->map_end = finite_size - 1;
    const fe_t er2 = fe_escape_r2(r->escape_radius_mode);
    for (int i = 1; i < finite_size; ++i) {
        if (fe_compare(fec_norm_squared(r->fast_orbit[i]), er2) > 0) { r->map_end = i; break; }
    }
    r->orbit_size = finite_size;
    r->scaled = (sc_t *)malloc(sizeof(sc_t) * (size_t)finite_size);
    r->real_double = (double *)malloc(sizeof(double) * (size_t)finite_size);
    r->imag_double = (double *)malloc(sizeof(double) * (size_t)finite_size);
    if (!r->scaled || !r->real_double || !r->imag_double) { fo_set_error(error, cap, "out of memory"); goto done; }
    for (int i = 0; i < finite_size; ++i) {
        r->scaled[i] = sc_from_fe(r->fast_orbit[i].re, r->fast_orbit[i].im);
        r->real_double[i] = ldexp(r->scaled[i].re, r->scaled[i].e);
        r->imag_double[i] = ldexp(r->scaled[i].im, r->scaled[i].e);
    }
    if (formula != 0) {
        stabilize_alternate_cycle(r);
        r->map_end = finite_size - 1;
        for (int i = 1; i < finite_size; ++i) {
            if (fe_compare(fec_norm_squared(r->fast_orbit[i]), er2) > 0) { r->map_end = i; break; }
        }
    }
    ok = 1;
done:
    mpfr_clears(cx, cy, vz, vr, zr, zi, nr, ni, tmp, norm, margin, pr, pi, ar, ai,
                dre, dim, dmag, rre, rim, are, aim, bre, bim, anorm, bnorm, dist, mult, tol, vlog, (mpfr_ptr)0);
    return ok;
}

/* stabilize_alternate_reference_cycle, renderer.cpp:9691-9788: a settled
 * projected cycle (three matching periods within 1e-12 relative, period
 * <= 64, found in the first 1024 entries) is repeated to the orbit end. */
static void stabilize_alternate_cycle(fo_reference *r) {
    const size_t count = (size_t)r->orbit_size;
    if (count < 6) return;
    size_t last_finite = count;
    while (last_finite > 0) {
        const size_t index = last_finite - 1;
        if (isfinite(r->real_double[index]) && isfinite(r->imag_double[index])) break;
        --last_finite;
    }
    if (last_finite < 6) return;
    const size_t scan_limit = last_finite - 1 < 1024 ? last_finite - 1 : 1024;
    const int max_period = (int)((scan_limit + 1) / 4) < 64 ? (int)((scan_limit + 1) / 4) : 64;
    const double tolerance = 1.0e-12;
    const double r2 = mode_r2_double(r->escape_radius_mode);
    for (int period = 1; period <= max_period; ++period) {
        for (size_t end = (size_t)(4 * period - 1); end <= scan_limit; ++end) {
            const size_t cycle_start = end - (size_t)(3 * period) + 1U;
            int matches = 1;
            for (int offset = 0; offset < period && matches; ++offset) {
                const size_t a = cycle_start + (size_t)offset, b = a + (size_t)period, c = b + (size_t)period;
                const double ar = r->real_double[a], ai = r->imag_double[a];
                const double br = r->real_double[b], bi = r->imag_double[b];
                const double cr = r->real_double[c], ci = r->imag_double[c];
                const double ra = hypot(ar, ai), rb = hypot(br, bi), rc = hypot(cr, ci);
                if (!isfinite(ra) || !isfinite(rb) || !isfinite(rc) || ra * ra >= r2 || rb * rb >= r2
                    || rc * rc >= r2) {
                    matches = 0;
                    break;
                }
                double scale = rb > rc ? rb : rc;
                scale = ra > scale ? ra : scale;
                scale = 1.0 > scale ? 1.0 : scale;
                const double d1 = hypot(br - ar, bi - ai);
                const double d2 = hypot(cr - br, ci - bi);
                if (!(d1 <= tolerance * scale && d2 <= tolerance * scale)) matches = 0;
            }
            if (!matches) continue;
            for (size_t target = cycle_start; target < count; ++target) {
                const size_t source = cycle_start + (target - cycle_start) % (size_t)period;
                r->fast_orbit[target] = r->fast_orbit[source];
                r->scaled[target] = r->scaled[source];
                r->real_double[target] = r->real_double[source];
                r->imag_double[target] = r->imag_double[source];
                r->margins[target] = r->margins[source];
            }
            return;
        }
    }
}

/* ---- image series (renderer.cpp:11162-11334) --------------------------- */
static fec_t series_evaluate(const fec_t *coefficients, int count, fec_t dc) { /* 11162-11174 */
    if (count <= 1) return fec_make(FE_ZERO, FE_ZERO);
    fec_t result = coefficients[count - 1];
    for (int index = count - 1; index > 1; --index)
        result = fec_add(fec_mul(result, dc), coefficients[index - 1]);
    return fec_mul(result, dc);
}

static int series_probe_is_safe(fec_t reference, fec_t exact, fec_t approximate,
                                fe_t minimum_scale_squared, fe_t tolerance_squared, int mode) {
    /* renderer.cpp:11176-11204 */
    const fec_t error = fec_sub(approximate, exact);
    fe_t scale = fec_norm_squared(exact);
    if (fe_compare(scale, minimum_scale_squared) < 0) scale = minimum_scale_squared;
    const fe_t allowed = fe_mul(scale, tolerance_squared);
    if (fe_compare(fec_norm_squared(error), allowed) > 0) return 0;
    const int exact_inside = fe_compare(fec_escape_margin_with_delta(reference, exact, mode), FE_ZERO) <= 0;
    const int approx_inside = fe_compare(fec_escape_margin_with_delta(reference, approximate, mode), FE_ZERO) <= 0;
    return exact_inside && approx_inside;
}

/* build_image_series, renderer.cpp:11206-11334 (Mandelbrot and Julia; for
 * Julia the pixel offset is the initial perturbation and c is fixed). */
static int build_image_series(fo_reference *r, const fec_t *orbit, int orbit_size) {
    const int julia = r->formula == 1;
    r->series_enabled = 0;
    r->series_order = 0;
    r->series_iteration = 1;
    r->series_radius_squared = FE_ZERO;
    free(r->series_coefficients);
    r->series_coefficients = NULL;
    r->series_count = 0;
    if (r->formula != 0 && !julia) return 1;
    if (orbit_size < 16 || r->requested_max_iter < 8) return 1;
    int order = r->requested_series_order;
    order = order < 8 ? 8 : (order > 32 ? 32 : order);
    int map_end = r->map_end;
    if (map_end < 2) map_end = 2;
    if (map_end > orbit_size - 1) map_end = orbit_size - 1;
    const int last_candidate = map_end - 1 > 2 ? map_end - 1 : 2;
    const fe_t viewport_radius = r->input_radius;
    if (fe_zero(viewport_radius) || !fe_finite(viewport_radius)) return 1;
    const fe_t probe_radius = fe_muld(viewport_radius, 1.5);
    const fe_t diagonal = fe_muld(probe_radius, 0.7071067811865476);
    const fe_t inner = fe_muld(probe_radius, 0.47);
    fec_t probes[12] = {
        {probe_radius, FE_ZERO}, {fe_neg(probe_radius), FE_ZERO},
        {FE_ZERO, probe_radius}, {FE_ZERO, fe_neg(probe_radius)},
        {diagonal, diagonal}, {fe_neg(diagonal), diagonal},
        {diagonal, fe_neg(diagonal)}, {fe_neg(diagonal), fe_neg(diagonal)},
        {inner, fe_muld(inner, 0.63)}, {fe_neg(inner), fe_muld(inner, 0.63)},
        {fe_muld(inner, 0.63), inner}, {fe_muld(inner, 0.63), fe_neg(inner)},
    };
    const int n = order + 1;
    fec_t *coef = (fec_t *)calloc((size_t)n, sizeof(fec_t));
    fec_t *next = (fec_t *)calloc((size_t)n, sizeof(fec_t));
    fec_t *best = (fec_t *)calloc((size_t)n, sizeof(fec_t));
    if (!coef || !next || !best) { free(coef); free(next); free(best); return 0; }
    fec_t exact[12];
    memset(exact, 0, sizeof(exact));
    if (julia) {
        coef[1] = fec_make(fe_from_parts(1.0, 0), FE_ZERO);
        for (int p = 0; p < 12; ++p) exact[p] = probes[p];
    }
    int best_iteration = 1, have_best = 0;
    const fe_t tolerance_squared = fe_from_parts(ldexp(1.0, -38), 0);
    const fe_t minimum_scale_squared = fe_sqr(probe_radius);
    const fe_t two = fe_from_parts(2.0, 0);
    const fec_t one = {fe_from_parts(1.0, 0), FE_ZERO};
    for (int iteration = 0; iteration < last_candidate; ++iteration) {
        const fec_t reference = orbit[iteration];
        for (int t = 0; t < n; ++t) next[t] = fec_make(FE_ZERO, FE_ZERO);
        for (int term = 1; term <= order; ++term) {
            fec_t value = fec_mulf(fec_mul(reference, coef[term]), two);
            for (int left = 1; left < term; ++left)
                value = fec_add(value, fec_mul(coef[left], coef[term - left]));
            if (!julia && term == 1) value = fec_add(value, one);
            next[term] = value;
        }
        fec_t *swap = coef; coef = next; next = swap;
        int safe = 1;
        for (int p = 0; p < 12; ++p) {
            const fec_t dc = probes[p];
            const fec_t exact_next = fec_add(fec_mulf(fec_mul(reference, exact[p]), two),
                                             fec_add(fec_mul(exact[p], exact[p]),
                                                     julia ? fec_make(FE_ZERO, FE_ZERO) : dc));
            exact[p] = exact_next;
            if (iteration + 1 >= order) {
                const fec_t approximate = series_evaluate(coef, n, dc);
                if (!series_probe_is_safe(orbit[iteration + 1], exact_next, approximate,
                                          minimum_scale_squared, tolerance_squared,
                                          r->escape_radius_mode)) { safe = 0; break; }
            }
        }
        if (!safe) break;
        if (iteration + 1 >= order) {
            best_iteration = iteration + 1;
            memcpy(best, coef, sizeof(fec_t) * (size_t)n);
            have_best = 1;
        }
    }
    if (best_iteration > 1 && have_best) {
        r->series_enabled = 1;
        r->series_order = order;
        r->series_iteration = best_iteration;
        r->series_radius_squared = fe_sqr(probe_radius);
        r->series_coefficients = (sc_t *)malloc(sizeof(sc_t) * (size_t)n);
        if (!r->series_coefficients) { free(coef); free(next); free(best); return 0; }
        for (int t = 0; t < n; ++t) r->series_coefficients[t] = sc_from_fe(best[t].re, best[t].im);
        r->series_count = n;
    }
    free(coef); free(next); free(best);
    return 1;
}

/* evaluate_image_series_with_derivative, renderer.cpp:8611-8639. */
static sc_t evaluate_series(const fo_reference *r, sc_t dc, sc_t *derivative, int want_derivative) {
    const int n = r->series_count;
    if (!r->series_enabled || n <= 1) { *derivative = SC_ZERO; return SC_ZERO; }
    sc_t value = r->series_coefficients[n - 1];
    sc_t pd = SC_ZERO;
    for (int index = n - 1; index > 1; --index) {
        if (want_derivative) pd = sc_add(sc_mul(pd, dc), value);
        value = sc_add(sc_mul(value, dc), r->series_coefficients[index - 1]);
    }
    if (want_derivative) *derivative = sc_add(sc_mul(pd, dc), value);
    return sc_mul(value, dc);
}

/* ---- BLA builders (renderer.cpp:11528-11769) --------------------------- */
static bla_builder_t merge_bla(const bla_builder_t *y, const bla_builder_t *x, fe_t input_radius) {
    /* renderer.cpp:11528-11586 */
    const fe_t x_a = fe_sqrt(fec_norm_squared(x->A));
    const fe_t x_b = fe_sqrt(fec_norm_squared(x->B));
    fe_t radius = fe_sqrt(x->r2);
    const fe_t remaining = fe_sub(fe_sqrt(y->r2), fe_mul(x_b, input_radius));
    if (fe_compare(remaining, FE_ZERO) > 0 && fe_compare(x_a, FE_ZERO) > 0)
        radius = fe_min(radius, fe_div(remaining, x_a));
    else
        radius = FE_ZERO;
    const fe_t two = fe_from_parts(2.0, 0), three = fe_from_parts(3.0, 0);
    const fec_t x_a2 = fec_mul(x->A, x->A);
    const fec_t x_ab = fec_mul(x->A, x->B);
    const fec_t x_b2 = fec_mul(x->B, x->B);
    const fec_t x_a3 = fec_mul(x_a2, x->A);
    const fec_t x_a2b = fec_mul(x_a2, x->B);
    const fec_t x_ab2 = fec_mul(x->A, x_b2);
    const fec_t x_b3 = fec_mul(x_b2, x->B);
    bla_builder_t s;
    s.A = fec_mul(y->A, x->A);
    s.B = fec_add(fec_mul(y->A, x->B), y->B);
    s.C = fec_add(fec_mul(y->A, x->C), fec_mul(y->C, x_a2));
    s.D = fec_add(fec_add(fec_mul(y->A, x->D), fec_mul(y->C, fec_mulf(x_ab, two))), fec_mul(y->D, x->A));
    s.E = fec_add(fec_add(fec_mul(y->A, x->E), fec_mul(y->C, x_b2)), fec_add(fec_mul(y->D, x->B), y->E));
    s.F = fec_add(fec_add(fec_mul(y->A, x->F), fec_mul(y->C, fec_mulf(fec_mul(x->A, x->C), two))),
                  fec_mul(y->F, x_a3));
    s.G = fec_add(
        fec_add(fec_add(fec_mul(y->A, x->G),
                        fec_mul(y->C, fec_mulf(fec_add(fec_mul(x->A, x->D), fec_mul(x->B, x->C)), two))),
                fec_add(fec_mul(y->D, x->C), fec_mul(y->F, fec_mulf(x_a2b, three)))),
        fec_mul(y->G, x_a2));
    s.H = fec_add(
        fec_add(fec_add(fec_mul(y->A, x->H),
                        fec_mul(y->C, fec_mulf(fec_add(fec_mul(x->A, x->E), fec_mul(x->B, x->D)), two))),
                fec_add(fec_mul(y->D, x->D), fec_mul(y->F, fec_mulf(x_ab2, three)))),
        fec_add(fec_mul(y->G, fec_mulf(x_ab, two)), fec_mul(y->H, x->A)));
    s.I = fec_add(
        fec_add(fec_add(fec_mul(y->A, x->I), fec_mul(y->C, fec_mulf(fec_mul(x->B, x->E), two))),
                fec_add(fec_mul(y->D, x->E), fec_mul(y->F, x_b3))),
        fec_add(fec_mul(y->G, x_b2), fec_add(fec_mul(y->H, x->B), y->I)));
    s.r2 = fe_sqr(radius);
    s.length = x->length + y->length;
    return s;
}

static linear_builder_t merge_linear_bla(const linear_builder_t *y, const linear_builder_t *x,
                                         fe_t input_radius) { /* renderer.cpp:11588-11610 */
    const fe_t x_a = fe_sqrt(fec_norm_squared(x->A));
    const fe_t x_b = fe_sqrt(fec_norm_squared(x->B));
    fe_t radius = fe_sqrt(x->r2);
    const fe_t remaining = fe_sub(fe_sqrt(y->r2), fe_mul(x_b, input_radius));
    if (fe_compare(remaining, FE_ZERO) > 0 && fe_compare(x_a, FE_ZERO) > 0)
        radius = fe_min(radius, fe_div(remaining, x_a));
    else
        radius = FE_ZERO;
    linear_builder_t s;
    s.A = fec_mul(y->A, x->A);
    s.B = fec_add(fec_mul(y->A, x->B), y->B);
    s.r2 = fe_sqr(radius);
    s.length = x->length + y->length;
    return s;
}

/* compact_bla_step, renderer.cpp:8645-8671. */
static cubic_step_t compact_cubic(const bla_builder_t *b) {
    const int finite = fe_finite(b->r2) && fec_finite(b->A) && fec_finite(b->B) && fec_finite(b->C)
        && fec_finite(b->D) && fec_finite(b->E) && fec_finite(b->F) && fec_finite(b->G)
        && fec_finite(b->H) && fec_finite(b->I);
    cubic_step_t c;
    const fec_t *k[9] = {&b->A, &b->B, &b->C, &b->D, &b->E, &b->F, &b->G, &b->H, &b->I};
    for (int i = 0; i < 9; ++i) c.c[i] = sc_from_fe(k[i]->re, k[i]->im);
    c.r2 = finite ? b->r2 : FE_ZERO;
    c.length = b->length;
    return c;
}

/* compact_linear_bla_step, renderer.cpp:8673-8684. */
static linear_step_t compact_linear(const linear_builder_t *b) {
    const int finite = fe_finite(b->r2) && fec_finite(b->A) && fec_finite(b->B);
    linear_step_t c;
    c.A = sc_from_fe(b->A.re, b->A.im);
    c.B = sc_from_fe(b->B.re, b->B.im);
    c.r2 = finite ? b->r2 : FE_ZERO;
    c.length = b->length;
    return c;
}

/* build_linear_levels lambda, renderer.cpp:11712-11757. */
static int build_linear_levels(fo_reference *r, int base_count, fe_t composition_radius, int tier) {
    const fe_t tolerance = fe_from_parts(ldexp(1.0, r->bla_epsilon_exponent), 0);
    const fe_t two = fe_from_parts(2.0, 0);
    int level_count = 1;
    size_t sizes[64];
    sizes[0] = (size_t)base_count;
    for (int level = 1; level < 64; ++level) {
        if ((1ULL << level) > (unsigned long long)MAX_SAFE_LINEAR_BLA_LENGTH) break;
        const size_t current = (sizes[level - 1] + 1) / 2;
        if (current == 0) break;
        sizes[level] = current;
        level_count = level + 1;
    }
    size_t total = 0;
    r->linear_offset[tier] = (size_t *)malloc(sizeof(size_t) * (size_t)level_count);
    r->linear_size[tier] = (size_t *)malloc(sizeof(size_t) * (size_t)level_count);
    for (int l = 0; l < level_count; ++l) {
        r->linear_offset[tier][l] = total;
        r->linear_size[tier][l] = sizes[l];
        total += sizes[l];
    }
    linear_builder_t *builder = (linear_builder_t *)malloc(sizeof(linear_builder_t) * total);
    r->linear[tier] = (linear_step_t *)malloc(sizeof(linear_step_t) * total);
    if (!builder || !r->linear[tier]) { free(builder); return 0; }
    for (int start = 1; start <= base_count; ++start) {
        const fec_t reference = r->fast_orbit[start];
        const fe_t radius = fe_mul(tolerance, fe_sqrt(fec_norm_squared(reference)));
        linear_builder_t s;
        s.A = fec_mulf(reference, two);
        s.B = fec_make(fe_from_parts(1.0, 0), FE_ZERO);
        s.r2 = fe_sqr(radius);
        s.length = 1;
        builder[start - 1] = s;
    }
    for (int level = 1; level < level_count; ++level) {
        const linear_builder_t *prev = builder + r->linear_offset[tier][level - 1];
        linear_builder_t *cur = builder + r->linear_offset[tier][level];
        const size_t prev_size = sizes[level - 1];
        for (size_t index = 0; index < sizes[level]; ++index) {
            const size_t first = index * 2;
            cur[index] = first + 1 < prev_size
                ? merge_linear_bla(&prev[first + 1], &prev[first], composition_radius)
                : prev[first];
        }
    }
    for (size_t i = 0; i < total; ++i) r->linear[tier][i] = compact_linear(&builder[i]);
    r->linear_levels[tier] = level_count;
    free(builder);
    return 1;
}

/* build_bla, Mandelbrot, renderer.cpp:11612-11769. */
static int build_bla(fo_reference *r) {
    const int max_iter = r->fast_orbit_size - 1;
    int map_end = r->map_end;
    if (map_end < 0) map_end = 0;
    if (map_end > max_iter) map_end = max_iter;
    int base_count = (max_iter < map_end ? max_iter : map_end) - 1;
    if (base_count < 0) base_count = 0;
    if (!build_image_series(r, r->fast_orbit, r->fast_orbit_size)) return 0;
    r->cubic_levels = 0;
    r->linear_levels[0] = r->linear_levels[1] = 0;
    if (base_count == 0) return 1;
    const fe_t tolerance = fe_from_parts(ldexp(1.0, r->bla_epsilon_exponent), 0);
    const fe_t two = fe_from_parts(2.0, 0);
    size_t sizes[64];
    int level_count = 1;
    sizes[0] = (size_t)base_count;
    for (int level = 1; level < 64; ++level) {
        if ((1ULL << level) > (unsigned long long)MAX_SAFE_BLA_LENGTH) break;
        const size_t current = sizes[level - 1] / 2;
        if (current == 0) break;
        sizes[level] = current;
        level_count = level + 1;
    }
    size_t total = 0;
    r->cubic_offset = (size_t *)malloc(sizeof(size_t) * (size_t)level_count);
    r->cubic_size = (size_t *)malloc(sizeof(size_t) * (size_t)level_count);
    for (int l = 0; l < level_count; ++l) {
        r->cubic_offset[l] = total;
        r->cubic_size[l] = sizes[l];
        total += sizes[l];
    }
    bla_builder_t *builder = (bla_builder_t *)malloc(sizeof(bla_builder_t) * total);
    r->cubic = (cubic_step_t *)malloc(sizeof(cubic_step_t) * total);
    if (!builder || !r->cubic) { free(builder); return 0; }
    const fec_t zero = fec_make(FE_ZERO, FE_ZERO);
    const fec_t one = fec_make(fe_from_parts(1.0, 0), FE_ZERO);
    for (int start = 1; start <= base_count; ++start) {
        const fec_t reference = r->fast_orbit[start];
        const fe_t radius = fe_mul(tolerance, fe_sqrt(fec_norm_squared(reference)));
        bla_builder_t s;
        s.A = fec_mulf(reference, two);
        s.B = one; s.C = one;
        s.D = zero; s.E = zero; s.F = zero; s.G = zero; s.H = zero; s.I = zero;
        s.r2 = fe_sqr(radius);
        s.length = 1;
        builder[start - 1] = s;
    }
    for (int level = 1; level < level_count; ++level) {
        const bla_builder_t *prev = builder + r->cubic_offset[level - 1];
        bla_builder_t *cur = builder + r->cubic_offset[level];
        for (size_t index = 0; index < sizes[level]; ++index)
            cur[index] = merge_bla(&prev[index * 2 + 1], &prev[index * 2], r->input_radius);
    }
    for (size_t i = 0; i < total; ++i) r->cubic[i] = compact_cubic(&builder[i]);
    r->cubic_levels = level_count;
    free(builder);
    r->deep_input_radius = fe_muld(r->input_radius, 1.0e-20);
    if (!build_linear_levels(r, base_count, r->input_radius, 0)) return 0;
    if (!build_linear_levels(r, base_count, r->deep_input_radius, 1)) return 0;
    return 1;
}

/* create_reference_context, Mandelbrot, renderer.cpp:13619-13702. */
static int build_alternate_bla(fo_reference *r);

fo_reference *fo_reference_create(const char *x, const char *y, const char *zoom, int max_iter,
                                  int precision_bits, int series_order, int escape_radius_mode,
                                  int coordinate_mode, char *error, int cap) {
    return fo_reference_create_formula(x, y, zoom, max_iter, precision_bits, series_order, escape_radius_mode,
                                       coordinate_mode, 0, "0", "0", error, cap);
}

fo_reference *fo_reference_create_formula(const char *x, const char *y, const char *zoom, int max_iter,
                                          int precision_bits, int series_order, int escape_radius_mode,
                                          int coordinate_mode, int formula, const char *julia_re,
                                          const char *julia_im, char *error, int cap) {
    return fo_reference_create_tuned(x, y, zoom, max_iter, precision_bits, series_order, escape_radius_mode,
                                     coordinate_mode, formula, julia_re, julia_im, -38, error, cap);
}

fo_reference *fo_reference_create_tuned(const char *x, const char *y, const char *zoom, int max_iter,
                                        int precision_bits, int series_order, int escape_radius_mode,
                                        int coordinate_mode, int formula, const char *julia_re,
                                        const char *julia_im, int bla_epsilon_exponent, char *error, int cap) {
    if (!x || !y || !zoom || max_iter <= 0 || precision_bits < 128 || formula < 0 || formula > 3 || !julia_re
        || !julia_im || bla_epsilon_exponent < -60 || bla_epsilon_exponent > -20) {
        fo_set_error(error, cap, "invalid reference configuration");
        return NULL;
    }
    fe_t parsed;
    if (!parse_zoom_fe(zoom, precision_bits, &parsed)) {
        fo_set_error(error, cap, "invalid deep Mandelbrot zoom");
        return NULL;
    }
    fo_reference *r = (fo_reference *)calloc(1, sizeof(fo_reference));
    if (!r) return NULL;
    r->escape_radius_mode = escape_radius_mode;
    r->coordinate_mode = coordinate_mode;
    r->bla_epsilon_exponent = bla_epsilon_exponent;
    r->requested_series_order = series_order < 8 ? 8 : (series_order > 32 ? 32 : series_order);
    r->requested_max_iter = max_iter;
    r->formula = formula;
    if (!make_orbit(r, x, y, zoom, max_iter, precision_bits, julia_re, julia_im, error, cap)
        || !(formula == 0 ? build_bla(r)
                          : build_image_series(r, r->fast_orbit, r->fast_orbit_size) && build_alternate_bla(r))) {
        fo_reference_destroy(r);
        return NULL;
    }
    return r;
}

/* retarget_bla_radius, renderer.cpp:11776-11800. */
static fe_t retarget_radius(fe_t x_r2, fe_t y_r2, sc_t xa, sc_t xb, fe_t input_radius) {
    const sn_t na = sc_norm_squared(xa), nb = sc_norm_squared(xb);
    const fe_t x_a = fe_sqrt(fe_make(na.m, na.e));
    const fe_t x_b = fe_sqrt(fe_make(nb.m, nb.e));
    fe_t radius = fe_sqrt(x_r2);
    const fe_t remaining = fe_sub(fe_sqrt(y_r2), fe_mul(x_b, input_radius));
    if (fe_compare(remaining, FE_ZERO) > 0 && fe_compare(x_a, FE_ZERO) > 0)
        radius = fe_min(radius, fe_div(remaining, x_a));
    else
        radius = FE_ZERO;
    radius = fe_muld(radius, 1.0 - ldexp(1.0, -40));
    return fe_sqr(radius);
}

static void copy_linear_tier(fo_reference *dst, const fo_reference *src, int tier_dst) {
    const int levels = src->linear_levels[0];
    size_t total = 0;
    for (int l = 0; l < levels; ++l) total += src->linear_size[0][l];
    dst->linear_levels[tier_dst] = levels;
    dst->linear_offset[tier_dst] = (size_t *)malloc(sizeof(size_t) * (size_t)(levels ? levels : 1));
    dst->linear_size[tier_dst] = (size_t *)malloc(sizeof(size_t) * (size_t)(levels ? levels : 1));
    dst->linear[tier_dst] = (linear_step_t *)malloc(sizeof(linear_step_t) * (total ? total : 1));
    memcpy(dst->linear_offset[tier_dst], src->linear_offset[0], sizeof(size_t) * (size_t)levels);
    memcpy(dst->linear_size[tier_dst], src->linear_size[0], sizeof(size_t) * (size_t)levels);
    memcpy(dst->linear[tier_dst], src->linear[0], sizeof(linear_step_t) * total);
}

static void retarget_linear(fo_reference *r, int tier, fe_t input_radius) { /* 11822-11845 */
    for (int level = 1; level < r->linear_levels[tier]; ++level) {
        const linear_step_t *prev = r->linear[tier] + r->linear_offset[tier][level - 1];
        linear_step_t *cur = r->linear[tier] + r->linear_offset[tier][level];
        const size_t prev_size = r->linear_size[tier][level - 1];
        for (size_t index = 0; index < r->linear_size[tier][level]; ++index) {
            const size_t first = index * 2;
            if (first + 1 >= prev_size) { cur[index].r2 = prev[first].r2; continue; }
            cur[index].r2 = retarget_radius(prev[first].r2, prev[first + 1].r2, prev[first].A,
                                            prev[first].B, input_radius);
        }
    }
}

/* clone_reference_context + build_retargeted_bla, renderer.cpp:13704-13768,
 * 11802-11868. */
fo_reference *fo_reference_clone(const fo_reference *s, const char *zoom, char *error, int cap) {
    if (!s || !zoom || !s->fast_orbit) { fo_set_error(error, cap, "invalid clone"); return NULL; }
    fe_t parsed;
    if (!parse_zoom_fe(zoom, s->precision_bits, &parsed)) {
        fo_set_error(error, cap, "invalid deep Mandelbrot zoom");
        return NULL;
    }
    fo_reference *r = (fo_reference *)calloc(1, sizeof(fo_reference));
    if (!r) return NULL;
    r->formula = s->formula;
    r->parameter = s->parameter;
    r->orbit_size = s->orbit_size;
    r->scaled = (sc_t *)malloc(sizeof(sc_t) * (size_t)s->orbit_size);
    r->real_double = (double *)malloc(sizeof(double) * (size_t)s->orbit_size);
    r->imag_double = (double *)malloc(sizeof(double) * (size_t)s->orbit_size);
    memcpy(r->scaled, s->scaled, sizeof(sc_t) * (size_t)s->orbit_size);
    memcpy(r->real_double, s->real_double, sizeof(double) * (size_t)s->orbit_size);
    memcpy(r->imag_double, s->imag_double, sizeof(double) * (size_t)s->orbit_size);
    r->map_end = s->map_end;
    r->requested_max_iter = s->requested_max_iter;
    r->requested_series_order = s->requested_series_order;
    r->escape_radius_mode = s->escape_radius_mode;
    r->coordinate_mode = s->coordinate_mode;
    r->bla_epsilon_exponent = s->bla_epsilon_exponent;
    r->precision_bits = s->precision_bits;
    r->input_radius = fe_div(fe_from_parts(viewport_height_factor(r->coordinate_mode), 0), parsed);
    if (s->formula != 0) {
        /* renderer.cpp:13739-13760: alternates rebuild series and matrices */
        r->fast_orbit_size = s->fast_orbit_size;
        r->fast_orbit = (fec_t *)malloc(sizeof(fec_t) * (size_t)s->fast_orbit_size);
        r->margins = (fe_t *)malloc(sizeof(fe_t) * (size_t)s->orbit_size);
        if (!r->fast_orbit || !r->margins) { fo_reference_destroy(r); return NULL; }
        memcpy(r->fast_orbit, s->fast_orbit, sizeof(fec_t) * (size_t)s->fast_orbit_size);
        memcpy(r->margins, s->margins, sizeof(fe_t) * (size_t)s->orbit_size);
        if (!build_image_series(r, r->fast_orbit, r->fast_orbit_size) || !build_alternate_bla(r)) {
            fo_reference_destroy(r);
            return NULL;
        }
        return r;
    }
    if (!build_image_series(r, s->fast_orbit, s->fast_orbit_size)) { fo_reference_destroy(r); return NULL; }
    /* cubic levels copied, then radii retargeted (renderer.cpp:11802-11820) */
    r->cubic_levels = s->cubic_levels;
    size_t total = 0;
    for (int l = 0; l < s->cubic_levels; ++l) total += s->cubic_size[l];
    r->cubic_offset = (size_t *)malloc(sizeof(size_t) * (size_t)(s->cubic_levels ? s->cubic_levels : 1));
    r->cubic_size = (size_t *)malloc(sizeof(size_t) * (size_t)(s->cubic_levels ? s->cubic_levels : 1));
    r->cubic = (cubic_step_t *)malloc(sizeof(cubic_step_t) * (total ? total : 1));
    memcpy(r->cubic_offset, s->cubic_offset, sizeof(size_t) * (size_t)s->cubic_levels);
    memcpy(r->cubic_size, s->cubic_size, sizeof(size_t) * (size_t)s->cubic_levels);
    memcpy(r->cubic, s->cubic, sizeof(cubic_step_t) * total);
    for (int level = 1; level < r->cubic_levels; ++level) {
        const cubic_step_t *prev = r->cubic + r->cubic_offset[level - 1];
        cubic_step_t *cur = r->cubic + r->cubic_offset[level];
        for (size_t index = 0; index < r->cubic_size[level]; ++index) {
            const cubic_step_t *xs = &prev[index * 2];
            const cubic_step_t *ys = &prev[index * 2 + 1];
            cur[index].r2 = retarget_radius(xs->r2, ys->r2, xs->c[0], xs->c[1], r->input_radius);
        }
    }
    copy_linear_tier(r, s, 0);
    retarget_linear(r, 0, r->input_radius);
    r->deep_input_radius = fe_muld(r->input_radius, 1.0e-20);
    copy_linear_tier(r, s, 1);
    retarget_linear(r, 1, r->deep_input_radius);
    return r;
}

int fo_reference_info(const fo_reference *r, int64_t *v, int capacity) {
    if (!r || !v || capacity < 9) return -1;
    v[0] = r->orbit_size; v[1] = r->map_end; v[2] = r->series_enabled;
    v[3] = r->series_iteration; v[4] = r->series_order; v[5] = r->cubic_levels;
    v[6] = r->linear_levels[0]; v[7] = r->linear_levels[1];
    v[8] = r->cubic_levels ? (int64_t)r->cubic_size[0] : 0;
    return 9;
}

int fo_reference_orbit(const fo_reference *r, double *re, double *im, int32_t *e, int capacity) {
    if (!r) return -1;
    const int n = r->orbit_size < capacity ? r->orbit_size : capacity;
    for (int i = 0; i < n; ++i) { re[i] = r->scaled[i].re; im[i] = r->scaled[i].im; e[i] = r->scaled[i].e; }
    return n;
}

/* ---- deep render (renderer.cpp:12450-13583) ---------------------------- */
enum { /* RenderStats field order, renderer.cpp:4615-4643 */
    ST_PIXELS, ST_LOGICAL, ST_BLA_BLOCKS, ST_LINEAR_BLOCKS, ST_CUBIC_BLOCKS, ST_EXACT,
    ST_REPLAY, ST_BLA_RETRIES, ST_CYCLE_INSIDE, ST_TAIL_PIXELS, ST_BLA_DISABLED,
    ST_TAIL_STEPS, ST_MAX_TAIL_STEPS, ST_TAIL_REBASES, ST_TAIL_FALLBACKS, ST_MAX_PIXEL_ITER,
    ST_SERIES_PIXELS, ST_SERIES_JUMPS, ST_GLITCH, ST_UNRESOLVED, ST_DEADLINE, ST_SECONDARY,
    ST_RENDER_NS, ST_HISTOGRAM
};

typedef struct {
    const fo_reference *r;
    int max_iter, mode, strict, allow_recovery;
    double bias;
    int disable_bla, cycle_enabled, strict_cycle, cycle_min_iter, cycle_margin;
    int use_deep_linear, max_bla_length, max_linear_length, approximation_order;
    int series_available;
    fe_t pixel_spacing;
    int budget_enabled;
    struct timespec deadline;
    atomic_int *budget_exceeded;
    FractalRenderPlanes *planes;
} render_ctx;

static int budget_expired(const render_ctx *c, unsigned *ticks) { /* renderer.cpp:11878-11890 */
    if (!c->budget_enabled) return 0;
    if (((*ticks)++ & 255U) != 0U) return 0;
    if (atomic_load_explicit(c->budget_exceeded, memory_order_relaxed)) return 1;
    struct timespec now;
    clock_gettime(CLOCK_MONOTONIC, &now);
    if (now.tv_sec > c->deadline.tv_sec
        || (now.tv_sec == c->deadline.tv_sec && now.tv_nsec >= c->deadline.tv_nsec)) {
        atomic_store_explicit(c->budget_exceeded, 1, memory_order_relaxed);
        return 1;
    }
    return 0;
}

static void record_bla_length(uint64_t *st, int length) { /* renderer.cpp:4767-4776 */
    if (!st || length <= 0) return;
    int bucket = 0;
    unsigned v = (unsigned)length;
    while (v > 1U && bucket < 15) { v >>= 1U; ++bucket; }
    ++st[ST_HISTOGRAM + bucket];
}

#define STAT(field) do { if (st) ++st[field]; } while (0)
#define STATN(field, n) do { if (st) st[field] += (uint64_t)(n); } while (0)

/* BlaLevels::lookup / lookup_linear, renderer.cpp:8816-8877. */
static int highest_level(int max_length) { return max_length > 0 ? 31 - __builtin_clz((unsigned)max_length) : 0; }

static const cubic_step_t *lookup_cubic(const fo_reference *r, int start, fe_t dn, int max_length) {
    if (start <= 0 || max_length <= 0 || !fe_finite(dn)) return NULL;
    const int base_count = r->cubic_levels ? (int)r->cubic_size[0] : 0;
    if (start > base_count) return NULL;
    int h = highest_level(max_length);
    if (h > r->cubic_levels - 1) h = r->cubic_levels - 1;
    const int offset = start - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= (int)r->cubic_size[level]) continue;
        const cubic_step_t *c = &r->cubic[r->cubic_offset[level] + (size_t)index];
        if (c->length > 1 && fe_finite(c->r2) && fe_compare(dn, c->r2) < 0) return c;
    }
    return NULL;
}

static const linear_step_t *lookup_linear(const fo_reference *r, int start, fe_t dn, int max_length, int deep) {
    const int tier = deep ? 1 : 0;
    if (start <= 0 || max_length <= 0 || !fe_finite(dn)) return NULL;
    const int base_count = r->linear_levels[tier] ? (int)r->linear_size[tier][0] : 0;
    if (start > base_count) return NULL;
    int h = highest_level(max_length);
    if (h > r->linear_levels[tier] - 1) h = r->linear_levels[tier] - 1;
    const int offset = start - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= (int)r->linear_size[tier][level]) continue;
        const linear_step_t *c = &r->linear[tier][r->linear_offset[tier][level] + (size_t)index];
        if (c->length > 1 && fe_finite(c->r2) && fe_compare(dn, c->r2) < 0) return c;
    }
    return NULL;
}

/* apply_bla_series, renderer.cpp:8686-8731. */
static sc_t apply_bla_series(const cubic_step_t *s, sc_t d, sc_t c, int order) {
    const sc_t *k = s->c;
    if (order <= 1) return sc_add(sc_mul(k[0], d), sc_mul(k[1], c));
    if (order == 2) {
        const sc_t d_inner = sc_add(sc_mul(k[2], d), sc_mul(k[3], c));
        const sc_t c_inner = sc_add(k[1], sc_mul(k[4], c));
        return sc_add(sc_mul(d, sc_add(k[0], d_inner)), sc_mul(c, c_inner));
    }
    const sc_t d_cubic = sc_add(sc_add(sc_mul(k[5], d), sc_mul(k[6], c)), k[2]);
    const sc_t c_cubic = sc_add(k[3], sc_mul(k[7], c));
    const sc_t d_inner = sc_add(k[0], sc_add(sc_mul(d, d_cubic), sc_mul(c, c_cubic)));
    const sc_t c_inner = sc_add(k[1], sc_mul(c, sc_add(k[4], sc_mul(k[8], c))));
    return sc_add(sc_mul(d, d_inner), sc_mul(c, c_inner));
}

/* apply_bla_series_derivative, renderer.cpp:8733-8795. */
static sc_t apply_bla_series_derivative(const cubic_step_t *s, sc_t d, sc_t c, sc_t der, int order) {
    const sc_t *k = s->c;
    if (order <= 1) return sc_add(sc_mul(k[0], der), k[1]);
    sc_t pd = k[0], pp = k[1];
    if (order >= 2) {
        pd = sc_add(pd, sc_add(sc_double(sc_mul(k[2], d)), sc_mul(k[3], c)));
        pp = sc_add(pp, sc_add(sc_mul(k[3], d), sc_double(sc_mul(k[4], c))));
    }
    if (order >= 3) {
        const sc_t d2 = sc_mul(d, d);
        const sc_t c2 = sc_mul(c, c);
        pd = sc_add(pd, sc_add(sc_mul(sc_add(sc_double(sc_mul(k[5], d)), sc_mul(k[5], d)), d),
                               sc_add(sc_double(sc_mul(k[6], sc_mul(d, c))), sc_mul(k[7], c2))));
        pp = sc_add(pp, sc_add(sc_mul(k[6], d2),
                               sc_add(sc_double(sc_mul(k[7], sc_mul(d, c))),
                                      sc_mul(sc_add(sc_double(sc_mul(k[8], c)), sc_mul(k[8], c)), c))));
    }
    return sc_add(sc_mul(pd, der), pp);
}

static const sc_t SC_ONE = {1.0, 0.0, 0};

/* render_scaled_double_tail, renderer.cpp:12450-12669.  Returns 1 when the
 * tail gave up (pathological), 0 otherwise; *unresolved / *deadline mirror
 * the reference's out-parameters. */
static int double_tail(const render_ctx *c, float *out, size_t index, sc_t dc, int *iteration,
                       int *reference_index, sc_t delta, sc_t *derivative, unsigned *ticks,
                       int *deadline, int *unresolved, uint64_t *st, sn_t initial_norm) {
    const fo_reference *r = c->r;
    FractalRenderPlanes *planes = c->planes;
    const int max_tail_rebases = 64, max_tail_steps = 65536;
    const double dc_re = sc_re_double(dc);
    const double dc_im = ldexp(dc.im, dc.e);
    const double escape_squared = mode_r2_double(c->mode);
    const double safe_escape_squared = escape_squared + 1.0e-7;
    double d_re = sc_re_double(delta);
    double d_im = ldexp(delta.im, delta.e);
    *out = encode_iteration(c->max_iter, c->bias);
    double tortoise_re = 0.0, tortoise_im = 0.0;
    int cycle_power = 1, cycle_length = 0, tail_steps = 0, tail_iterations = 0, tail_rebases = 0;
    sn_t previous_norm = initial_norm;
    int cycle_ready = 0;
    const int cycle_on = c->cycle_enabled;
    while (*iteration < c->max_iter && *reference_index >= 0 && *reference_index + 1 < r->orbit_size) {
        if (budget_expired(c, ticks)) {
            *out = NAN; *deadline = 1; *unresolved = 1;
            STAT(ST_DEADLINE); STAT(ST_UNRESOLVED);
            return 0;
        }
        const double z_re = r->real_double[*reference_index];
        const double z_im = r->imag_double[*reference_index];
        if (planes) {
            sc_t prior_total = sc_make(z_re + d_re, z_im + d_im, 0);
            sc_normalize(&prior_total);
            *derivative = sc_add(sc_double(sc_mul(prior_total, *derivative)), SC_ONE);
        }
        const double lin_re = 2.0 * (z_re * d_re - z_im * d_im);
        const double lin_im = 2.0 * (z_re * d_im + z_im * d_re);
        const double sq_re = d_re * d_re - d_im * d_im;
        const double sq_im = 2.0 * d_re * d_im;
        d_re = lin_re + sq_re + dc_re;
        d_im = lin_im + sq_im + dc_im;
        ++*reference_index;
        ++*iteration;
        if (++tail_iterations > max_tail_steps) {
            STAT(ST_TAIL_FALLBACKS); STAT(ST_UNRESOLVED);
            *unresolved = 1;
            *out = c->strict ? NAN : encode_iteration(*iteration, c->bias);
            return 1;
        }
        STAT(ST_LOGICAL); STAT(ST_EXACT);
        const double t_re = r->real_double[*reference_index] + d_re;
        const double t_im = r->imag_double[*reference_index] + d_im;
        const double m2 = t_re * t_re + t_im * t_im;
        if (!isfinite(m2)) {
            *unresolved = 1;
            STAT(ST_GLITCH); STAT(ST_UNRESOLVED);
            *out = c->strict ? NAN : encode_iteration(*iteration, c->bias);
            return 0;
        }
        if (m2 > escape_squared) {
            const double magnitude = sqrt(m2 > safe_escape_squared ? m2 : safe_escape_squared);
            const double q = log(log(magnitude)) / (double)LOG_TWO_L;
            *out = (float)(((long double)*iteration - (long double)q) - (long double)c->bias);
            if (planes) {
                sc_t escaped_total = sc_make(t_re, t_im, 0);
                sc_normalize(&escaped_total);
                store_planes_escape(planes, index, kalles_raw_iteration(*iteration), escaped_total,
                                    scaled_norm_from_double(m2), previous_norm, *derivative,
                                    c->pixel_spacing);
            }
            return 0;
        }
        previous_norm = scaled_norm_from_double(m2);
        const double dm2 = d_re * d_re + d_im * d_im;
        ++tail_steps;
        if (cycle_on && tail_steps >= 64 && (tail_steps & 31) == 0) {
            if (!cycle_ready) {
                tortoise_re = t_re; tortoise_im = t_im;
                cycle_power = 1; cycle_length = 0; cycle_ready = 1;
            } else {
                const double cdr = t_re - tortoise_re, cdi = t_im - tortoise_im;
                const double cd2 = cdr * cdr + cdi * cdi;
                const int cmin = c->strict_cycle ? 2048 : 512;
                const double ctol = c->strict_cycle ? 1.0e-24 : 1.0e-18;
                if (*iteration >= cmin && m2 < 3.0 && cd2 <= ctol * (1.0 < m2 ? m2 : 1.0)) {
                    clear_planes_pixel(planes, index, c->max_iter);
                    return 0;
                }
                ++cycle_length;
                if (cycle_length >= cycle_power) {
                    tortoise_re = t_re; tortoise_im = t_im;
                    cycle_power = cycle_power * 2 < (1 << 20) ? cycle_power * 2 : (1 << 20);
                    cycle_length = 0;
                }
            }
        }
        if (m2 < dm2) {
            d_re = t_re; d_im = t_im;
            *reference_index = 0;
            tail_steps = 0;
            cycle_ready = 0;
            STAT(ST_TAIL_REBASES);
            if (++tail_rebases > max_tail_rebases) {
                STAT(ST_TAIL_FALLBACKS); STAT(ST_UNRESOLVED);
                *unresolved = 1;
                *out = c->strict ? NAN : encode_iteration(*iteration, c->bias);
                clear_planes_pixel(planes, index, c->max_iter);
                return 1;
            }
        }
    }
    if (*iteration < c->max_iter) {
        *unresolved = 1;
        STAT(ST_UNRESOLVED);
        *out = c->strict ? NAN : encode_iteration(*iteration, c->bias);
        clear_planes_pixel(planes, index, c->max_iter);
    }
    return 0;
}

/* The exact scaled recurrence after the reference escaped or after too
 * many rebases, renderer.cpp:13044-13128 and 13155-13250.  Returns 1 if the
 * pixel ended by deadline. */
static int exact_tail(const render_ctx *c, float *out, size_t index, sc_t dc, int *iteration,
                      sc_t total, sn_t total_norm, sc_t *derivative, unsigned *ticks, int *deadline,
                      int *unresolved, uint64_t *st) {
    const fo_reference *r = c->r;
    const sc_t parameter = sc_add(r->scaled[1], dc);
    sc_t tortoise = SC_ZERO;
    int power = 1, length = 0, ready = 0, escaped = 0;
    while (*iteration < c->max_iter) {
        if (budget_expired(c, ticks)) {
            *out = NAN; *deadline = 1; *unresolved = 1;
            STAT(ST_DEADLINE); STAT(ST_UNRESOLVED);
            break;
        }
        const sn_t prior_norm = total_norm;
        if (c->planes) *derivative = sc_add(sc_double(sc_mul(total, *derivative)), SC_ONE);
        total = sc_add(sc_mul(total, total), parameter);
        ++*iteration;
        STAT(ST_LOGICAL); STAT(ST_EXACT);
        total_norm = sc_norm_squared(total);
        if (sc_outside_escape(total_norm, c->mode)) {
            *out = smooth_escape_scaled(*iteration, total_norm, c->bias);
            store_planes_escape(c->planes, index, kalles_raw_iteration(*iteration), total, total_norm,
                                prior_norm, *derivative, c->pixel_spacing);
            escaped = 1;
            break;
        }
        if (c->cycle_enabled && *iteration >= c->cycle_min_iter && (*iteration & 31) == 0
            && sc_compare_norm(total_norm, (sn_t){0.75, 2}) < 0) {
            if (!ready) {
                tortoise = total; power = 1; length = 0; ready = 1;
            } else {
                const sn_t distance = sc_norm_squared(sc_sub(total, tortoise));
                const int scale_exponent = total_norm.e > 0 ? total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale_exponent - c->cycle_margin) {
                    *out = encode_iteration(c->max_iter, c->bias);
                    escaped = 1;
                    break;
                }
                ++length;
                if (length >= power) {
                    tortoise = total;
                    power = power * 2 < (1 << 20) ? power * 2 : (1 << 20);
                    length = 0;
                }
            }
        }
    }
    if (*deadline) return 1;
    if (!escaped) *out = encode_iteration(c->max_iter, c->bias);
    return 0;
}

/* One pixel of render_bla_impl, renderer.cpp:12817-13544. */
static void render_pixel(const render_ctx *c, float *output, size_t index, sc_t dc, uint64_t *st) {
    const fo_reference *r = c->r;
    FractalRenderPlanes *planes = c->planes;
    const int max_iter = c->max_iter;
    STAT(ST_PIXELS);
    clear_planes_pixel(planes, index, max_iter);
    sc_t delta = dc;
    sc_t derivative = SC_ONE;
    int reference_index = 1, iteration = 1;
    if (c->series_available
        && sc_compare_norm(sc_norm_squared(dc), (sn_t){r->series_radius_squared.m, r->series_radius_squared.e}) <= 0) {
        delta = evaluate_series(r, dc, &derivative, planes != NULL);
        reference_index = r->series_iteration;
        iteration = r->series_iteration;
        STAT(ST_SERIES_PIXELS);
        STATN(ST_SERIES_JUMPS, iteration);
    }
    int escaped = 0, have_total = 0;
    sc_t total = SC_ZERO;
    sn_t total_norm = {0.0, 0}, previous_total_norm = {0.0, 0};
    sc_t tortoise = SC_ZERO;
    int cycle_power = 1, cycle_length = 0, cycle_ready = 0;
    int pixel_disable_bla = 0, rebases = 0;
    unsigned ticks = 0;
    int deadline = 0, unresolved = 0;
    float *out = &output[index];
    const fe_t glitch_scale = fe_from_parts(1.0e-7, 0);

    while (iteration < max_iter) {
        if (budget_expired(c, &ticks)) {
            *out = NAN; deadline = 1; unresolved = 1;
            STAT(ST_DEADLINE); STAT(ST_UNRESOLVED);
            break;
        }
        if (reference_index < 0 || reference_index >= r->orbit_size) {
            *out = encode_iteration(iteration, c->bias);
            store_planes_escape(planes, index, kalles_raw_iteration(iteration), total, total_norm,
                                previous_total_norm, derivative, c->pixel_spacing);
            escaped = 1;
            break;
        }
        const sc_t z = r->scaled[reference_index];
        fe_t margin = FE_ZERO;
        if (!have_total) {
            total = sc_add(z, delta);
            margin = sc_escape_margin_with_delta(z, delta, c->mode);
            total_norm = sc_norm_with_margin(margin, c->mode);
        }
        const int had_total = have_total;
        have_total = 0;
        /* glitch test, renderer.cpp:12931-12988 */
        const sn_t zn = sc_norm_squared(z);
        const fe_t reference_norm = fe_make(zn.m, zn.e);
        if (reference_norm.m != 0.0 && fe_finite(reference_norm)
            && fe_compare(fe_make(total_norm.m, total_norm.e), fe_mul(reference_norm, glitch_scale)) < 0) {
            STAT(ST_GLITCH);
            if (c->allow_recovery && sc_finite(total) && ++rebases <= 64) {
                delta = total;
                reference_index = 0;
                have_total = 0;
                continue;
            }
            *out = NAN;
            unresolved = 1;
            STAT(ST_UNRESOLVED);
            break;
        }
        /* escape test, renderer.cpp:12989-13006 */
        if (had_total) margin = sc_escape_margin_with_delta(z, delta, c->mode);
        if (sc_outside_escape(total_norm, c->mode) || fe_compare(margin, FE_ZERO) > 0) {
            *out = smooth_escape_scaled(iteration, total_norm, c->bias);
            store_planes_escape(planes, index, kalles_raw_iteration(iteration), total, total_norm,
                                previous_total_norm, derivative, c->pixel_spacing);
            escaped = 1;
            break;
        }
        /* cycle detection, renderer.cpp:13016-13042 */
        if (c->cycle_enabled && iteration >= c->cycle_min_iter && (iteration & 31) == 0
            && sc_compare_norm(total_norm, (sn_t){0.75, 2}) < 0) {
            if (!cycle_ready) {
                tortoise = total; cycle_power = 1; cycle_length = 0; cycle_ready = 1;
            } else {
                const sn_t distance = sc_norm_squared(sc_sub(total, tortoise));
                const int scale_exponent = total_norm.e > 0 ? total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale_exponent - c->cycle_margin) {
                    STAT(ST_CYCLE_INSIDE);
                    *out = encode_iteration(max_iter, c->bias);
                    escaped = 1;
                    break;
                }
                ++cycle_length;
                if (cycle_length >= cycle_power) {
                    tortoise = total;
                    cycle_power = cycle_power * 2 < (1 << 20) ? cycle_power * 2 : (1 << 20);
                    cycle_length = 0;
                }
            }
        }
        /* rebase, renderer.cpp:13047-13153 */
        const sn_t delta_norm = sc_norm_squared(delta);
        if (sc_compare_norm(total_norm, delta_norm) < 0) {
            if (++rebases > 64) {
                exact_tail(c, out, index, dc, &iteration, total, total_norm, &derivative, &ticks,
                           &deadline, &unresolved, st);
                if (!deadline) escaped = 1;
                break;
            }
            delta = total;
            reference_index = 0;
            continue;
        }
        /* reference escaped: direct scaled recurrence, renderer.cpp:13155-13250 */
        if (r->map_end > 1 && reference_index >= r->map_end) {
            exact_tail(c, out, index, dc, &iteration, total, total_norm, &derivative, &ticks,
                       &deadline, &unresolved, st);
            if (!deadline) escaped = 1;
            break;
        }
        /* BLA selection, renderer.cpp:13252-13273 */
        const fe_t dn = fe_make(delta_norm.m, delta_norm.e);
        const int remaining = max_iter - iteration;
        const int effective_order = c->approximation_order >= 3 && delta_norm.e < -115
            ? (delta_norm.e < -160 ? 1 : 2) : c->approximation_order;
        const linear_step_t *linear = NULL;
        const cubic_step_t *cubic = NULL;
        if (!c->disable_bla && !pixel_disable_bla) {
            if (effective_order <= 1)
                linear = lookup_linear(r, reference_index, dn,
                                       c->max_linear_length < remaining ? c->max_linear_length : remaining,
                                       c->use_deep_linear);
            else
                cubic = lookup_cubic(r, reference_index, dn,
                                     c->max_bla_length < remaining ? c->max_bla_length : remaining);
        }
        const int map_length = linear ? linear->length : (cubic ? cubic->length : 0);
        /* double tail, renderer.cpp:13283-13362 */
        if (!c->disable_bla && !pixel_disable_bla
            && (map_length <= 0 || reference_index + map_length > max_iter)
            && delta_norm.e > -90) {
            STAT(ST_TAIL_PIXELS);
            const int tail_start = iteration;
            const int pathological = double_tail(c, out, index, dc, &iteration, &reference_index, delta,
                                                 &derivative, &ticks, &deadline, &unresolved, st,
                                                 total_norm);
            if (deadline || unresolved) break;
            if (st) {
                const uint64_t steps = (uint64_t)(iteration - tail_start > 0 ? iteration - tail_start : 0);
                st[ST_TAIL_STEPS] += steps;
                if (steps > st[ST_MAX_TAIL_STEPS]) st[ST_MAX_TAIL_STEPS] = steps;
            }
            if (pathological) {
                delta = dc;
                derivative = SC_ONE;
                reference_index = 1;
                iteration = 1;
                have_total = 0;
                pixel_disable_bla = 1;
                continue;
            }
            escaped = 1;
            break;
        }
        /* BLA jump, renderer.cpp:13363-13514 */
        if (map_length > 0 && reference_index + map_length <= max_iter) {
            STAT(ST_BLA_BLOCKS);
            if (linear) STAT(ST_LINEAR_BLOCKS); else STAT(ST_CUBIC_BLOCKS);
            STAT(ST_SERIES_JUMPS);
            record_bla_length(st, map_length);
            const sc_t previous_delta = delta, previous_derivative = derivative;
            const int previous_reference = reference_index, previous_iteration = iteration;
            const sc_t input_delta = delta;
            delta = linear ? sc_add(sc_mul(linear->A, input_delta), sc_mul(linear->B, dc))
                           : apply_bla_series(cubic, input_delta, dc, effective_order);
            if (planes)
                derivative = linear ? sc_add(sc_mul(linear->A, previous_derivative), linear->B)
                                    : apply_bla_series_derivative(cubic, input_delta, dc,
                                                                  previous_derivative, effective_order);
            reference_index += map_length;
            iteration += map_length;
            const sc_t zr = r->scaled[reference_index];
            const sc_t endpoint = sc_add(zr, delta);
            const sn_t endpoint_norm = sc_norm_with_margin(sc_escape_margin_with_delta(zr, delta, c->mode), c->mode);
            if (!sc_finite(delta) || !sc_finite(endpoint)) {
                delta = previous_delta; derivative = previous_derivative;
                reference_index = previous_reference; iteration = previous_iteration;
                STAT(ST_BLA_RETRIES);
                if (!pixel_disable_bla) STAT(ST_BLA_DISABLED);
                pixel_disable_bla = 1;
                have_total = 0;
                continue;
            }
            if (sc_compare_norm(endpoint_norm, (sn_t){0.75, 2}) >= 0) {
                delta = previous_delta; derivative = previous_derivative;
                reference_index = previous_reference; iteration = previous_iteration;
                sc_t replay_total = SC_ZERO;
                sn_t replay_norm = {0.0, 0};
                int retry = 0;
                for (int replay = 0; replay < map_length && iteration < max_iter; ++replay) {
                    if (budget_expired(c, &ticks)) {
                        *out = NAN; deadline = 1; unresolved = 1;
                        STAT(ST_DEADLINE); STAT(ST_UNRESOLVED);
                        break;
                    }
                    const sc_t reference = r->scaled[reference_index];
                    if (planes) {
                        const sc_t prior_total = sc_add(reference, delta);
                        derivative = sc_add(sc_double(sc_mul(prior_total, derivative)), SC_ONE);
                    }
                    delta = sc_add(sc_double(sc_mul(reference, delta)), sc_add(sc_mul(delta, delta), dc));
                    ++reference_index;
                    ++iteration;
                    STAT(ST_REPLAY); STAT(ST_LOGICAL);
                    const sc_t zn2 = r->scaled[reference_index];
                    replay_total = sc_add(zn2, delta);
                    const fe_t rm = sc_escape_margin_with_delta(zn2, delta, c->mode);
                    replay_norm = sc_norm_with_margin(rm, c->mode);
                    if (!sc_finite(delta) || !sc_finite(replay_total)) { retry = 1; break; }
                    if (sc_outside_escape(replay_norm, c->mode) || fe_compare(rm, FE_ZERO) > 0) {
                        *out = smooth_escape_scaled(iteration, replay_norm, c->bias);
                        store_planes_escape(planes, index, kalles_raw_iteration(iteration), replay_total,
                                            replay_norm, total_norm, derivative, c->pixel_spacing);
                        escaped = 1;
                        break;
                    }
                }
                if (deadline || escaped) break;
                if (retry) {
                    delta = previous_delta; derivative = previous_derivative;
                    reference_index = previous_reference; iteration = previous_iteration;
                    STAT(ST_BLA_RETRIES);
                    if (!pixel_disable_bla) STAT(ST_BLA_DISABLED);
                    pixel_disable_bla = 1;
                    have_total = 0;
                    continue;
                }
                total = replay_total;
                total_norm = replay_norm;
                previous_total_norm = total_norm;
                have_total = 1;
            } else {
                total = endpoint;
                total_norm = endpoint_norm;
                previous_total_norm = total_norm;
                have_total = 1;
                STATN(ST_LOGICAL, map_length);
            }
            continue;
        }
        /* exact perturbation step, renderer.cpp:13516-13535 */
        const sn_t prior_norm = total_norm;
        if (planes) {
            const sc_t prior_total = sc_add(z, delta);
            derivative = sc_add(sc_double(sc_mul(prior_total, derivative)), SC_ONE);
        }
        delta = sc_add(sc_double(sc_mul(z, delta)), sc_add(sc_mul(delta, delta), dc));
        ++reference_index;
        ++iteration;
        previous_total_norm = prior_norm;
        STAT(ST_LOGICAL); STAT(ST_EXACT);
    }
    if (!escaped && !deadline && !unresolved) *out = encode_iteration(max_iter, c->bias);
    if (st && (uint64_t)(iteration > 0 ? iteration : 0) > st[ST_MAX_PIXEL_ITER])
        st[ST_MAX_PIXEL_ITER] = (uint64_t)(iteration > 0 ? iteration : 0);
}

typedef struct { int x, y, width, height; } fo_rect;

static int render_alternate(float *output, int width, int height, fe_t zoom, const fo_reference *r, int max_iter,
                            const FractalRenderOptions *o, uint64_t *stats, int threads, struct timespec started,
                            const sc_t *points, const fe_t *point_radius, const fo_rect *crop);

/* render_bla_impl's field loop (renderer.cpp:12672-13583) over a grid, or
 * over caller-supplied normalised dc offsets (width = count, height = 1)
 * when points is set.  With crop set, only the pixels of that rectangle of
 * the width x height grid are rendered, into a crop->width x crop->height
 * output: every per-pixel offset and every field-wide decision is the full
 * grid's. */
static int render_mandel_field(float *output, int width, int height, fe_t zoom, const fo_reference *r, int max_iter,
                             int series_order, int series_block, const FractalRenderOptions *options,
                             FractalRenderPlanes *planes, uint64_t *stats, int threads, const sc_t *points,
                             const fe_t *point_radius, struct timespec started, char *error, int cap,
                             const fo_rect *crop) {
    render_ctx c;
    memset(&c, 0, sizeof(c));
    c.r = r;
    c.max_iter = max_iter;
    c.mode = options->escape_radius_mode;
    c.strict = options->strict != 0;
    c.allow_recovery = options->allow_recovery != 0;
    c.bias = options->output_bias;
    c.planes = planes;
    /* viewport, renderer.cpp:12689-12705 */
    const double factor = viewport_height_factor(options->coordinate_mode);
    const fe_t inverse_zoom = fe_div(fe_from_parts(1.0, 0), zoom);
    const fe_t view_height = fe_muld(inverse_zoom, factor);
    const fe_t view_width = fe_muld(view_height, (double)width / (double)height);
    c.pixel_spacing = fe_div(view_height, fe_from_parts((double)height, 0));
    /* a point render uses its farthest point's radius (renderer.cpp:12701-12705) */
    const fe_t current_input_radius = points ? *point_radius : fe_muld(inverse_zoom, factor);
    const int covers = fe_compare(current_input_radius, r->input_radius) <= 0;
    c.disable_bla = !covers || options->disable_bla != 0;
    c.cycle_enabled = options->disable_cycle == 0 && (options->strict == 0 || options->strict_cycle != 0);
    c.strict_cycle = options->strict_cycle != 0;
    c.cycle_min_iter = c.strict_cycle ? 2048 : 512;
    c.cycle_margin = c.strict_cycle ? 78 : 60;
    if (options->time_budget_ms > 0) {
        c.budget_enabled = 1;
        clock_gettime(CLOCK_MONOTONIC, &c.deadline);
        long long ns = (long long)c.deadline.tv_nsec + (long long)options->time_budget_ms * 1000000LL;
        c.deadline.tv_sec += (time_t)(ns / 1000000000LL);
        c.deadline.tv_nsec = (long)(ns % 1000000000LL);
    }
    atomic_int exceeded;
    atomic_init(&exceeded, 0);
    c.budget_exceeded = &exceeded;
    c.use_deep_linear = fe_compare(current_input_radius, r->deep_input_radius) <= 0;
    const fe_t ultra = fe_muld(r->input_radius, 1.0e-40);
    const int use_ultra = fe_compare(current_input_radius, ultra) <= 0;
    int mbl = series_block < options->max_bla_length ? series_block : options->max_bla_length;
    c.max_bla_length = mbl < 2 ? 2 : (mbl > MAX_SAFE_BLA_LENGTH ? MAX_SAFE_BLA_LENGTH : mbl);
    const int linear_cap = use_ultra ? MAX_SAFE_LINEAR_BLA_LENGTH
        : (c.use_deep_linear ? MAX_SAFE_DEEP_LINEAR_BLA_LENGTH : MAX_SAFE_BLA_LENGTH);
    int mll = series_block < options->max_linear_bla_length ? series_block : options->max_linear_bla_length;
    c.max_linear_length = mll < 2 ? 2 : (mll > linear_cap ? linear_cap : mll);
    c.approximation_order = series_order < 1 ? 1 : (series_order > 3 ? 3 : series_order);
    c.series_available = r->series_enabled && r->series_order >= options->series_min_terms
        && r->series_order <= options->series_max_terms
        && fe_compare(current_input_radius, fe_sqrt(r->series_radius_squared)) <= 0
        && r->series_iteration < max_iter;
    /* per-row / per-column offsets, renderer.cpp:12764-12783 */
    fe_t *xo = (fe_t *)malloc(sizeof(fe_t) * (size_t)(points ? 1 : width));
    fe_t *yo = (fe_t *)malloc(sizeof(fe_t) * (size_t)(points ? 1 : height));
    if (!xo || !yo) { free(xo); free(yo); fo_set_error(error, cap, "out of memory"); return 1; }
    for (int py = 0; py < height && !points; ++py)
        yo[py] = fe_muld(view_height, -pixel_axis_offset(py, height, options->coordinate_mode) / (double)height);
    for (int px = 0; px < width && !points; ++px)
        xo[px] = fe_muld(view_width, pixel_axis_offset(px, width, options->coordinate_mode) / (double)width);
#ifdef _OPENMP
    if (threads > 0) { omp_set_dynamic(0); omp_set_num_threads(threads); }
    const int workers = omp_get_max_threads();
#else
    (void)threads;
    const int workers = 1;
#endif
    uint64_t *local = stats ? (uint64_t *)calloc((size_t)workers * FO_STATS_FIELDS, sizeof(uint64_t)) : NULL;
    const fo_rect area = crop ? *crop : (fo_rect){0, 0, width, height};
    const long long count = (long long)area.width * (long long)area.height;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (long long lp = 0; lp < count; ++lp) {
        const int py = area.y + (int)(lp / area.width);
        const int px = area.x + (int)(lp % area.width);
#ifdef _OPENMP
        uint64_t *st = local ? local + (size_t)omp_get_thread_num() * FO_STATS_FIELDS : NULL;
#else
        uint64_t *st = local;
#endif
        render_pixel(&c, output, (size_t)lp, points ? points[lp] : sc_from_fe(xo[px], yo[py]), st);
    }
    if (stats) {
        memset(stats, 0, sizeof(uint64_t) * FO_STATS_FIELDS);
        for (int w = 0; w < workers; ++w) {
            const uint64_t *s = local + (size_t)w * FO_STATS_FIELDS;
            for (int f = 0; f < FO_STATS_FIELDS; ++f) {
                if (f == ST_MAX_TAIL_STEPS || f == ST_MAX_PIXEL_ITER)
                    stats[f] = stats[f] > s[f] ? stats[f] : s[f];
                else
                    stats[f] += s[f];
            }
        }
        struct timespec ended;
        clock_gettime(CLOCK_MONOTONIC, &ended);
        stats[ST_RENDER_NS] = (uint64_t)((ended.tv_sec - started.tv_sec) * 1000000000LL
                                         + (ended.tv_nsec - started.tv_nsec));
    }
    free(local); free(xo); free(yo);
    return 0;
}

static int render_grid(float *output, int width, int height, const char *zoom_text, const fo_reference *r,
                       int max_iter, int series_order, int series_block, const FractalRenderOptions *options,
                       FractalRenderPlanes *planes, uint64_t *stats, int threads, char *error, int cap,
                       const fo_rect *crop) {
    struct timespec started;
    clock_gettime(CLOCK_MONOTONIC, &started);
    if (!output || !r || !options || width <= 0 || height <= 0) {
        fo_set_error(error, cap, "invalid render arguments");
        return 1;
    }
    if (crop && (crop->x < 0 || crop->y < 0 || crop->width <= 0 || crop->height <= 0
                 || crop->width > width - crop->x || crop->height > height - crop->y)) {
        fo_set_error(error, cap, "crop rectangle outside the field");
        return 1;
    }
    fe_t zoom;
    if (!parse_zoom_fe(zoom_text, r->precision_bits, &zoom)) {
        fo_set_error(error, cap, "invalid deep Mandelbrot zoom");
        return 1;
    }
    if (r->formula != 0) {
        if (planes) {
            fo_set_error(error, cap, "render metadata requires the native scalar backend for this formula");
            return 1;
        }
        if (render_alternate(output, width, height, zoom, r, max_iter, options, stats, threads, started, NULL,
                             NULL, crop)) {
            fo_set_error(error, cap, "out of memory");
            return 1;
        }
        return 0;
    }
    return render_mandel_field(output, width, height, zoom, r, max_iter, series_order, series_block, options, planes,
                             stats, threads, NULL, NULL, started, error, cap, crop);
}

int fo_render(float *output, int width, int height, const char *zoom_text, const fo_reference *r,
              int max_iter, int series_order, int series_block, const FractalRenderOptions *options,
              FractalRenderPlanes *planes, uint64_t *stats, int threads, char *error, int cap) {
    return render_grid(output, width, height, zoom_text, r, max_iter, series_order, series_block, options, planes,
                       stats, threads, error, cap, NULL);
}

int fo_render_crop(float *output, int width, int height, int crop_x, int crop_y, int crop_width, int crop_height,
                   const char *zoom_text, const fo_reference *r, int max_iter, int series_order, int series_block,
                   const FractalRenderOptions *options, FractalRenderPlanes *planes, uint64_t *stats, int threads,
                   char *error, int cap) {
    const fo_rect crop = {crop_x, crop_y, crop_width, crop_height};
    return render_grid(output, width, height, zoom_text, r, max_iter, series_order, series_block, options, planes,
                       stats, threads, error, cap, &crop);
}

/* fractal_render_points_impl, renderer.cpp:20019-20186, for every formula:
 * each point is checked finite and normalised, the largest |dc|^2 gives the
 * radius that replaces the viewport's in the BLA, series and tier
 * decisions.  The scalar backend's strict-mode MPFR repair of a
 * glitched point (renderer.cpp:12944-12966, at most 8192 points) is not
 * restated: glitched points stay NaN, as on the reference's GPU point path
 * (renderer.cpp:996-1010). */
int fo_render_points(float *output, int point_count, const char *zoom_text, const double *real_mantissa,
                     const double *imag_mantissa, const int32_t *exponents, const fo_reference *r, int max_iter,
                     int series_order, int series_block, const FractalRenderOptions *options, uint64_t *stats,
                     int threads, char *error, int cap) {
    struct timespec started;
    clock_gettime(CLOCK_MONOTONIC, &started);
    if (!output || !r || !options || point_count <= 0 || !real_mantissa || !imag_mantissa || !exponents) {
        fo_set_error(error, cap, "invalid native point-render arguments");
        return 1;
    }
    fe_t zoom;
    if (!parse_zoom_fe(zoom_text, r->precision_bits, &zoom)) {
        fo_set_error(error, cap, "invalid deep Mandelbrot zoom");
        return 1;
    }
    sc_t *points = (sc_t *)malloc(sizeof(sc_t) * (size_t)point_count);
    if (!points) { fo_set_error(error, cap, "out of memory"); return 1; }
    sn_t largest = {0.0, 0};
    for (int i = 0; i < point_count; ++i) {
        sc_t point = {real_mantissa[i], imag_mantissa[i], exponents[i]};
        if (!sc_finite(point)) {
            free(points);
            fo_set_error(error, cap, "non-finite point offset");
            return 1;
        }
        sc_normalize(&point);
        points[i] = point;
        const sn_t radius = sc_norm_squared(point);
        if (sc_compare_norm(radius, largest) > 0) largest = radius;
    }
    const fe_t point_radius = fe_sqrt((fe_t){largest.m, largest.e});
    if (r->formula != 0) { /* renderer.cpp:20105-20136 */
        const int failed = render_alternate(output, point_count, 1, zoom, r, max_iter, options, stats, threads,
                                            started, points, &point_radius, NULL);
        free(points);
        if (failed) fo_set_error(error, cap, "out of memory");
        return failed;
    }
    const int status = render_mandel_field(output, point_count, 1, zoom, r, max_iter, series_order, series_block,
                                         options, NULL, stats, threads, points, &point_radius, started, error, cap,
                                         NULL);
    free(points);
    return status;
}

/* ---- shallow direct render (renderer.cpp:10009-10426) -------------------- */

/* parse_classic_long_double + range checks (renderer.cpp:214-241, 7516-7530). */
static int parse_long_double(const char *text, long double *out) {
    if (!text) return 0;
    char *end = NULL;
    const long double v = strtold(text, &end);
    if (end == text) return 0;
    while (*end == ' ' || *end == '\t' || *end == '\n' || *end == '\r' || *end == '\f' || *end == '\v') ++end;
    if (*end != '\0' || !isfinite(v)) return 0;
    *out = v;
    return 1;
}

typedef struct {
    const double *x, *y;
    float *output;
    FractalRenderPlanes *planes;
    int width, height, max_iter, formula;
    double julia_re, julia_im, bias, escape2, safe_escape2, log_power;
    uint32_t flags;
    fe_t spacing;
} direct_ctx;

/* iterate_direct_formula_with_derivative, renderer.cpp:251-398, with each
 * product and sum rounded in the reference's order. */
static void direct_step(int formula, int derivative, double zr, double zi, double pr, double pi, double dr,
                        double di, double *nr, double *ni, double *ndr, double *ndi) {
    if (formula == 2) {  /* burning ship */
        const double ar = fabs(zr), ai = fabs(zi);
        const double rs = ar * ar, is = ai * ai, cross = 2.0 * ar * ai;
        *nr = (rs - is) + pr;
        *ni = cross + pi;
    } else if (formula == 3) {  /* tricorn */
        const double rs = zr * zr, is = zi * zi, cross = -2.0 * zr * zi;
        *nr = (rs - is) + pr;
        *ni = cross + pi;
    } else {  /* Mandelbrot and Julia share the map */
        *nr = zr * zr - zi * zi + pr;
        *ni = 2.0 * zr * zi + pi;
    }
    *ndr = 0.0;
    *ndi = 0.0;
    if (derivative) {  /* Mandelbrot (parameter plane, +1) and Julia (+0) */
        const double add_real = formula == 1 ? 0.0 : 1.0;
        *ndr = 2.0 * zr * dr - 2.0 * zi * di + add_real;
        *ndi = 2.0 * zr * di + 2.0 * zi * dr + 0.0;
    }
}

static void direct_inside(const direct_ctx *c, size_t index) {
    c->output[index] = encode_iteration(c->max_iter, c->bias);
    clear_planes_pixel(c->planes, index, c->max_iter);
}

static void direct_pixel(const direct_ctx *c, int px, int py) {
    const double cx = c->x[px], cy = c->y[py];
    const size_t index = (size_t)py * (size_t)c->width + (size_t)px;
    const int julia = c->formula == 1;
    const int derivative = c->planes && (c->flags & 2u) == 0 && (c->formula == 0 || julia);
    if (c->formula == 0) { /* cardioid / period-2 bulb, renderer.cpp:10087-10097 */
        const double q = (cx - 0.25) * (cx - 0.25) + cy * cy;
        const int in_cardioid = q * (q + cx - 0.25) <= 0.25 * cy * cy;
        const int in_bulb = (cx + 1.0) * (cx + 1.0) + cy * cy <= 0.0625;
        if (in_cardioid || in_bulb) { direct_inside(c, index); return; }
    }
    double zr = julia ? cx : 0.0, zi = julia ? cy : 0.0;
    const double pr = julia ? c->julia_re : cx, pi = julia ? c->julia_im : cy;
    double dr = julia ? 1.0 : 0.0, di = 0.0;
    double test1 = 0.0, test2 = 0.0;
    int iteration = 0;
    for (; iteration < c->max_iter; ++iteration) {
        double nr, ni, ndr, ndi;
        direct_step(c->formula, derivative, zr, zi, pr, pi, dr, di, &nr, &ni, &ndr, &ndi);
        zr = nr; zi = ni; dr = ndr; di = ndi;
        const double m2 = zr * zr + zi * zi;
        test2 = test1;
        test1 = m2;
        if (m2 > c->escape2 || !isfinite(m2)) {
            const double safe = isfinite(m2) ? (m2 < c->safe_escape2 ? c->safe_escape2 : m2) : DBL_MAX;
            const double magnitude = sqrt(safe);
            c->output[index] = (float)(((long double)(iteration + 1) - log(log(magnitude)) / c->log_power)
                                       - (long double)c->bias);
            if (c->planes) {
                FractalRenderPlanes *p = c->planes;
                p->orbit_iteration[index] = iteration;
                p->phase[index] = 0.0;
                if ((c->flags & 1u) == 0) {
                    double value = 0.0;
                    if (isfinite(zr) && isfinite(zi)) {
                        value = atan2(zi, zr) / 6.283185307179586476925286766559005768;
                        value -= floor(value);
                    }
                    p->phase[index] = value;
                }
                p->de_x[index] = 0.0;
                p->de_y[index] = 0.0;
                if (derivative && isfinite(zr) && isfinite(zi) && isfinite(dr) && isfinite(di)) {
                    const sc_t total = sc_from_fe(fe_from_parts(zr, 0), fe_from_parts(zi, 0));
                    const sc_t dz = sc_from_fe(fe_from_parts(dr, 0), fe_from_parts(di, 0));
                    (void)render_plane_de(total, dz, c->spacing, &p->de_x[index], &p->de_y[index]);
                }
                p->test1[index] = test1;
                p->test2[index] = test2;
            }
            return;
        }
    }
    direct_inside(c, index);
}

int fo_render_direct(float *output, int width, int height, const char *zoom, const char *x_center,
                     const char *y_center, int max_iter, int formula, double julia_real, double julia_imag,
                     double output_bias, int escape_radius_mode, int coordinate_mode, uint32_t plane_flags,
                     FractalRenderPlanes *planes, int threads, char *error, int cap) {
    long double zl = 0.0L, xl = 0.0L, yl = 0.0L;
    if (!output || width <= 0 || height <= 0 || max_iter <= 0 || formula < 0 || formula > 3
        || !parse_long_double(zoom, &zl) || !(zl > 0.0L) || !parse_long_double(x_center, &xl)
        || !parse_long_double(y_center, &yl)) {
        fo_set_error(error, cap, "invalid direct render arguments");
        return 1;
    }
    const long double l10 = log10l(zl);
    if (!isfinite(l10) || l10 < -300.0L || l10 > 9800.0L) {
        fo_set_error(error, cap, "Mandelbrot zoom is outside the supported range");
        return 1;
    }
    const double zoom_value = (double)zl, centre_re = (double)xl, centre_im = (double)yl;
    if (!isfinite(zoom_value) || zoom_value <= 0.0) {
        fo_set_error(error, cap, "direct native coordinates or zoom exceed double range");
        return 1;
    }
    const double height_span = viewport_height_factor(coordinate_mode) / zoom_value;
    const double width_span = height_span * (double)width / (double)height;
    if (!isfinite(height_span) || !isfinite(width_span)) {
        fo_set_error(error, cap, "direct native viewport is outside double range");
        return 1;
    }
    double *xs = malloc(sizeof(double) * (size_t)width);
    double *ys = malloc(sizeof(double) * (size_t)height);
    if (!xs || !ys) { free(xs); free(ys); fo_set_error(error, cap, "out of memory"); return 1; }
    for (int px = 0; px < width; ++px)
        xs[px] = centre_re + pixel_axis_offset(px, width, coordinate_mode) * width_span / (double)width;
    for (int py = 0; py < height; ++py)
        ys[py] = centre_im + -pixel_axis_offset(py, height, coordinate_mode) * height_span / (double)height;
    direct_ctx c;
    c.x = xs; c.y = ys; c.output = output; c.planes = planes;
    c.width = width; c.height = height; c.max_iter = max_iter; c.formula = formula;
    c.julia_re = julia_real; c.julia_im = julia_imag; c.bias = output_bias;
    c.escape2 = mode_r2_double(escape_radius_mode);
    c.safe_escape2 = c.escape2 + 1.0e-7;
    c.log_power = log(2.0);
    c.flags = plane_flags;
    c.spacing = fe_from_parts(height_span / (double)height, 0);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int py = 0; py < height; ++py)
        for (int px = 0; px < width; ++px) direct_pixel(&c, px, py);
    free(xs); free(ys);
    return 0;
}

/* ---- alternate formulas (renderer.cpp:11336-11530, 11917-12448) ---------- */

/* AlternateLinearBlaStep, renderer.cpp:8520-8525: d' = M d + P dc with
 * M = c[0..3] and P = c[4..7], both real 2x2 row-major. */
typedef struct alt_step { fe_t c[8]; fe_t r2; int length; } alt_step_t;

static inline fe_t fe_abs(fe_t v) { return fe_make(fabs(v.m), v.e); }

static fe_t alt_matrix_norm(const fe_t *c, int offset) { /* renderer.cpp:11340-11350 */
    fe_t sum = FE_ZERO;
    for (int k = 0; k < 4; ++k) sum = fe_add(sum, fe_sqr(c[offset + k]));
    return fe_sqrt(sum);
}

static alt_step_t merge_alt(const alt_step_t *y, const alt_step_t *x, fe_t input_parameter_radius) {
    /* merge_alternate_linear_bla, renderer.cpp:11352-11397 */
    const fe_t *yc = y->c, *xc = x->c;
    alt_step_t out;
    out.c[0] = fe_add(fe_mul(yc[0], xc[0]), fe_mul(yc[1], xc[2]));
    out.c[1] = fe_add(fe_mul(yc[0], xc[1]), fe_mul(yc[1], xc[3]));
    out.c[2] = fe_add(fe_mul(yc[2], xc[0]), fe_mul(yc[3], xc[2]));
    out.c[3] = fe_add(fe_mul(yc[2], xc[1]), fe_mul(yc[3], xc[3]));
    out.c[4] = fe_add(fe_add(fe_mul(yc[0], xc[4]), fe_mul(yc[1], xc[6])), yc[4]);
    out.c[5] = fe_add(fe_add(fe_mul(yc[0], xc[5]), fe_mul(yc[1], xc[7])), yc[5]);
    out.c[6] = fe_add(fe_add(fe_mul(yc[2], xc[4]), fe_mul(yc[3], xc[6])), yc[6]);
    out.c[7] = fe_add(fe_add(fe_mul(yc[2], xc[5]), fe_mul(yc[3], xc[7])), yc[7]);
    const fe_t x_state = alt_matrix_norm(xc, 0);
    const fe_t x_parameter = alt_matrix_norm(xc, 4);
    fe_t radius = fe_sqrt(x->r2);
    const fe_t remaining = fe_sub(fe_sqrt(y->r2), fe_mul(x_parameter, input_parameter_radius));
    if (fe_compare(remaining, FE_ZERO) > 0 && fe_compare(x_state, FE_ZERO) > 0)
        radius = fe_min(radius, fe_div(remaining, x_state));
    else
        radius = FE_ZERO;
    radius = fe_muld(radius, 1.0 - ldexp(1.0, -40));
    out.r2 = fe_sqr(radius);
    out.length = x->length + y->length;
    return out;
}

/* build_alternate_linear_bla, renderer.cpp:11424-11526. */
static int build_alternate_bla(fo_reference *r) {
    free(r->alt); free(r->alt_offset); free(r->alt_size);
    r->alt = NULL; r->alt_offset = NULL; r->alt_size = NULL;
    r->alt_levels = 0;
    r->alt_input_radius_squared = fe_sqr(r->input_radius);
    r->alt_start = r->formula == 1 ? 0 : 1;
    if (r->fast_orbit_size < 2) return 1;
    const int first = r->alt_start;
    const int last = r->map_end < r->fast_orbit_size - 1 ? r->map_end : r->fast_orbit_size - 1;
    const int base_count = last - first;
    if (base_count <= 0) return 1;
    size_t sizes[64], total = 0;
    int levels = 0;
    sizes[levels++] = (size_t)base_count;
    total += (size_t)base_count;
    for (int level = 1; level < 64; ++level) {
        if ((1ULL << level) > 4096ULL) break;  /* MAX_SAFE_LINEAR_BLA_LENGTH */
        const size_t current = (sizes[level - 1] + 1) / 2;
        if (current == 0) break;
        sizes[levels++] = current;
        total += current;
    }
    r->alt = (alt_step_t *)calloc(total, sizeof(alt_step_t));
    r->alt_offset = (size_t *)malloc(sizeof(size_t) * (size_t)levels);
    r->alt_size = (size_t *)malloc(sizeof(size_t) * (size_t)levels);
    if (!r->alt || !r->alt_offset || !r->alt_size) return 0;
    size_t offset = 0;
    for (int l = 0; l < levels; ++l) { r->alt_offset[l] = offset; r->alt_size[l] = sizes[l]; offset += sizes[l]; }
    r->alt_levels = levels;
    const fe_t tolerance = fe_from_parts(ldexp(1.0, r->bla_epsilon_exponent), 0);
    const fe_t two = fe_from_parts(2.0, 0), quarter = fe_from_parts(0.25, 0), one = fe_from_parts(1.0, 0);
    for (int k = 0; k < base_count; ++k) {
        const fec_t ref = r->fast_orbit[first + k];
        const fe_t re = ref.re, im = ref.im;
        alt_step_t *b = &r->alt[k];
        memset(b, 0, sizeof(*b));
        if (r->formula == 3) {
            b->c[0] = fe_mul(re, two);
            b->c[1] = fe_muld(im, -2.0);
            b->c[2] = fe_muld(im, -2.0);
            b->c[3] = fe_muld(re, -2.0);
        } else if (r->formula == 2) {
            const fe_t are = fe_abs(re), aim = fe_abs(im);
            const double rs = re.m < 0.0 ? -1.0 : 1.0, is = im.m < 0.0 ? -1.0 : 1.0;
            b->c[0] = fe_muld(are, 2.0 * rs);
            b->c[1] = fe_muld(aim, -2.0 * is);
            b->c[2] = fe_muld(aim, 2.0 * rs);
            b->c[3] = fe_muld(are, 2.0 * is);
        } else {
            b->c[0] = fe_mul(re, two);
            b->c[1] = fe_muld(im, -2.0);
            b->c[2] = fe_mul(im, two);
            b->c[3] = fe_mul(re, two);
        }
        if (r->formula != 1) {
            b->c[4] = one;
            b->c[7] = one;
        }
        const fe_t magnitude = fe_sqrt(fec_norm_squared(ref));
        fe_t radius = fe_mul(tolerance, fe_compare(magnitude, one) > 0 ? magnitude : one);
        if (r->formula == 2) {
            if (fe_zero(re) || fe_zero(im)) {
                radius = FE_ZERO;
            } else {
                radius = fe_min(radius, fe_mul(fe_abs(re), quarter));
                radius = fe_min(radius, fe_mul(fe_abs(im), quarter));
            }
        }
        b->r2 = fe_sqr(radius);
        b->length = 1;
    }
    for (int level = 1; level < levels; ++level) {
        const alt_step_t *prev = r->alt + r->alt_offset[level - 1];
        alt_step_t *cur = r->alt + r->alt_offset[level];
        const size_t prev_size = r->alt_size[level - 1];
        for (size_t index = 0; index < r->alt_size[level]; ++index) {
            const size_t f = index * 2;
            cur[index] = f + 1 < prev_size ? merge_alt(&prev[f + 1], &prev[f], r->input_radius) : prev[f];
        }
    }
    return 1;
}

/* AlternateLinearBlaLevels::lookup, renderer.cpp:8533-8565. */
static const alt_step_t *alt_lookup(const fo_reference *r, int start, fe_t dn, fe_t pn, int max_length) {
    if (r->alt_levels == 0 || start < r->alt_start || max_length <= 0 || !fe_finite(dn) || !fe_finite(pn))
        return NULL;
    const int offset = start - r->alt_start;
    if (offset < 0 || offset >= (int)r->alt_size[0]) return NULL;
    int h = 31 - __builtin_clz((unsigned)max_length);
    if (h > r->alt_levels - 1) h = r->alt_levels - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= (int)r->alt_size[level]) continue;
        const alt_step_t *c = &r->alt[r->alt_offset[level] + (size_t)index];
        if (c->length > 1 && fe_finite(c->r2) && fe_compare(dn, c->r2) < 0
            && fe_compare(pn, r->alt_input_radius_squared) <= 0)
            return c;
    }
    return NULL;
}

static sc_t alt_apply(const alt_step_t *s, sc_t delta, sc_t parameter) { /* renderer.cpp:11399-11416 */
    const fe_t dr = sc_part(delta, 0), di = sc_part(delta, 1);
    const fe_t pr = sc_part(parameter, 0), pi = sc_part(parameter, 1);
    const fe_t rr = fe_add(fe_add(fe_mul(s->c[0], dr), fe_mul(s->c[1], di)),
                           fe_add(fe_mul(s->c[4], pr), fe_mul(s->c[5], pi)));
    const fe_t ri = fe_add(fe_add(fe_mul(s->c[2], dr), fe_mul(s->c[3], di)),
                           fe_add(fe_mul(s->c[6], pr), fe_mul(s->c[7], pi)));
    return sc_from_fe(rr, ri);
}

/* scaled_formula_step, renderer.cpp:11917-11931. */
static sc_t alt_formula_step(int formula, sc_t v, sc_t parameter) {
    if (formula == 2) {
        const sc_t a = sc_make(fabs(v.re), fabs(v.im), v.e);
        return sc_add(sc_mul(a, a), parameter);
    }
    const sc_t sq = formula == 3 ? sc_mul(sc_conj(v), sc_conj(v)) : sc_mul(v, v);
    return sc_add(sq, parameter);
}

/* alternate_delta_step, renderer.cpp:11933-12034. */
static sc_t alt_delta_step(const fo_reference *r, int i, sc_t delta, sc_t pdelta) {
    const sc_t z = r->scaled[i];
    if (r->formula == 3) {
        const sc_t cz = sc_conj(z), cd = sc_conj(delta);
        return sc_add(sc_double(sc_mul(cz, cd)), sc_add(sc_mul(cd, cd), pdelta));
    }
    if (r->formula == 2) {
        const fe_t zr = sc_part(z, 0), zi = sc_part(z, 1);
        const fe_t dr = sc_part(delta, 0), di = sc_part(delta, 1);
        const fe_t azr = fe_abs(zr), azi = fe_abs(zi);
        const fe_t tr = fe_add(zr, dr), ti = fe_add(zi, di);
        const int crosses_real = (fe_compare(zr, FE_ZERO) >= 0) != (fe_compare(tr, FE_ZERO) >= 0);
        const int crosses_imag = (fe_compare(zi, FE_ZERO) >= 0) != (fe_compare(ti, FE_ZERO) >= 0);
        if (crosses_real || crosses_imag) {
            const fe_t atr = fe_abs(tr), ati = fe_abs(ti);
            const sc_t next = r->scaled[i + 1];
            const fe_t por = fe_sub(sc_part(r->parameter, 0), sc_part(next, 0));
            const fe_t poi = fe_sub(sc_part(r->parameter, 1), sc_part(next, 1));
            const fe_t nr = fe_add(fe_sub(fe_sqr(atr), fe_sqr(ati)), fe_add(por, sc_part(pdelta, 0)));
            const fe_t ni = fe_add(fe_muld(fe_mul(atr, ati), 2.0), fe_add(poi, sc_part(pdelta, 1)));
            return sc_from_fe(nr, ni);
        }
        const double rs = zr.m < 0.0 ? -1.0 : 1.0, is = zi.m < 0.0 ? -1.0 : 1.0;
        const fe_t rl = fe_sub(fe_muld(fe_mul(azr, dr), 2.0 * rs), fe_muld(fe_mul(azi, di), 2.0 * is));
        const fe_t il = fe_add(fe_muld(fe_mul(azr, di), 2.0 * is), fe_muld(fe_mul(azi, dr), 2.0 * rs));
        const fe_t sr = fe_sub(fe_sqr(dr), fe_sqr(di));
        const fe_t si = fe_muld(fe_mul(dr, di), 2.0 * rs * is);
        return sc_from_fe(fe_add(fe_add(rl, sr), sc_part(pdelta, 0)), fe_add(fe_add(il, si), sc_part(pdelta, 1)));
    }
    return sc_add(sc_double(sc_mul(z, delta)), sc_add(sc_mul(delta, delta), pdelta));
}

/* alternate_escape_margin via sc_escape_margin_with_reference_margin,
 * renderer.cpp:8435-8451, 12036-12053. */
static fe_t alt_margin(const fo_reference *r, int i, sc_t delta) {
    const sc_t z = r->scaled[i];
    const fe_t cross = fe_muld(fe_add(fe_mul(sc_part(z, 0), sc_part(delta, 0)),
                                      fe_mul(sc_part(z, 1), sc_part(delta, 1))), 2.0);
    const sn_t dn = sc_norm_squared(delta);
    return fe_add(fe_add(r->margins[i], cross), fe_make(dn.m, dn.e));
}

typedef struct {
    const fo_reference *r;
    float *output;
    int width, max_iter, mode, strict, cycle_enabled, series_available, bla_enabled, max_block;
    double bias;
    const fe_t *xo, *yo;
    const sc_t *points; /* point render: one normalised dc per pixel index, or NULL */
    int crop_x, crop_y, out_width; /* output holds the out_width-wide rectangle at (crop_x, crop_y) */
    int budget_enabled;
    struct timespec deadline;
    atomic_int *budget_exceeded;
} alt_ctx;

static void alt_pixel(const alt_ctx *c, int px, int py) {
    const fo_reference *r = c->r;
    const size_t index = (size_t)(py - c->crop_y) * (size_t)c->out_width + (size_t)(px - c->crop_x);
    const int julia = r->formula == 1;
    const int mode = c->mode;
    const sc_t dc = c->points ? c->points[(size_t)py * (size_t)c->width + (size_t)px]
                              : sc_from_fe(c->xo[px], c->yo[py]);
    const sc_t pdelta = julia ? SC_ZERO : dc;
    const sn_t pnorm = sc_norm_squared(pdelta);
    const fe_t pn = fe_make(pnorm.m, pnorm.e);
    sc_t delta = dc;
    int ref = julia ? 0 : 1, iteration = julia ? 0 : 1;
    if (c->series_available) {
        const sn_t dcn = sc_norm_squared(dc);
        const sn_t rad = {r->series_radius_squared.m, r->series_radius_squared.e};
        if (sc_compare_norm(dcn, rad) <= 0) {
            sc_t unused;
            delta = evaluate_series(r, dc, &unused, 0);
            ref = r->series_iteration;
            iteration = r->series_iteration;
        }
    }
    unsigned ticks = 0;
    sc_t total = sc_add(r->scaled[ref], delta);
    sn_t total_norm = sc_norm_with_margin(alt_margin(r, ref, delta), mode);
    (void)total;
    if (sc_outside_escape(total_norm, mode)) {
        c->output[index] = smooth_escape_scaled(iteration, total_norm, c->bias);
        return;
    }
    sc_t tortoise = SC_ZERO;
    int cycle_power = 1, cycle_length = 0, cycle_ready = 0, cycle_hits = 0;
    while (iteration < c->max_iter) {
        if (c->budget_enabled && (ticks++ & 255U) == 0U) {
            int expired = atomic_load(c->budget_exceeded);
            if (!expired) {
                struct timespec now;
                clock_gettime(CLOCK_MONOTONIC, &now);
                if (now.tv_sec > c->deadline.tv_sec
                    || (now.tv_sec == c->deadline.tv_sec && now.tv_nsec >= c->deadline.tv_nsec)) {
                    atomic_store(c->budget_exceeded, 1);
                    expired = 1;
                }
            }
            if (expired) { c->output[index] = NAN; return; }
        }
        const int next = ref + 1;
        if (next >= r->orbit_size) {
            const sc_t parameter = julia ? r->parameter : sc_add(r->parameter, dc);
            total = alt_formula_step(r->formula, total, parameter);
            ++iteration;
            total_norm = sc_norm_squared(total);
            if (sc_outside_escape(total_norm, mode)) {
                c->output[index] = smooth_escape_scaled(iteration, total_norm, c->bias);
                return;
            }
            continue;
        }
        const sn_t dn = sc_norm_squared(delta);
        const int remaining = c->max_iter - iteration;
        const alt_step_t *step = c->bla_enabled
            ? alt_lookup(r, ref, fe_make(dn.m, dn.e), pn, c->max_block < remaining ? c->max_block : remaining)
            : NULL;
        if (step && ref + step->length < r->orbit_size) {
            const int end = ref + step->length;
            const sc_t cand = alt_apply(step, delta, pdelta);
            const fe_t cm = alt_margin(r, end, cand);
            const sn_t cn = sc_norm_with_margin(cm, mode);
            const sc_t ct = sc_add(r->scaled[end], cand);
            const sn_t limit = {0.75, 2};
            const int bad = !sc_finite(cand) || !fe_finite(cm) || !sc_finite(ct) || fe_compare(cm, FE_ZERO) > 0
                || sc_compare_norm(cn, limit) >= 0;
            if (!bad) {
                delta = cand;
                ref = end;
                iteration += step->length;
                total = ct;
                total_norm = cn;
                continue;
            }
            for (int replay = 0; replay < step->length && iteration < c->max_iter; ++replay) {
                delta = alt_delta_step(r, ref, delta, pdelta);
                ++ref;
                ++iteration;
                if (!sc_finite(delta)) {
                    c->output[index] = c->strict ? NAN : encode_iteration(iteration, c->bias);
                    return;
                }
                const fe_t m = alt_margin(r, ref, delta);
                total_norm = sc_norm_with_margin(m, mode);
                total = sc_add(r->scaled[ref], delta);
                if (fe_compare(m, FE_ZERO) > 0 || sc_outside_escape(total_norm, mode)) {
                    c->output[index] = smooth_escape_scaled(iteration, total_norm, c->bias);
                    return;
                }
            }
            continue;
        }
        delta = alt_delta_step(r, ref, delta, pdelta);
        ref = next;
        ++iteration;
        if (!sc_finite(delta)) {
            c->output[index] = c->strict ? NAN : encode_iteration(iteration, c->bias);
            return;
        }
        const fe_t m = alt_margin(r, ref, delta);
        total_norm = sc_norm_with_margin(m, mode);
        total = sc_add(r->scaled[ref], delta);
        if (fe_compare(m, FE_ZERO) > 0 || sc_outside_escape(total_norm, mode)) {
            c->output[index] = smooth_escape_scaled(iteration, total_norm, c->bias);
            return;
        }
        const sn_t limit = {0.75, 2};
        if (c->cycle_enabled && iteration >= 2048 && (iteration & 31) == 0
            && sc_compare_norm(total_norm, limit) < 0) {
            if (!cycle_ready) {
                tortoise = total; cycle_power = 1; cycle_length = 0; cycle_hits = 0; cycle_ready = 1;
            } else {
                const sn_t distance = sc_norm_squared(sc_sub(total, tortoise));
                const int scale = total_norm.e > 0 ? total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale - 78) {
                    if (++cycle_hits >= 3) {
                        c->output[index] = encode_iteration(c->max_iter, c->bias);
                        return;
                    }
                } else {
                    cycle_hits = 0;
                }
                if (++cycle_length >= cycle_power) {
                    tortoise = total;
                    cycle_power = cycle_power * 2 < (1 << 20) ? cycle_power * 2 : (1 << 20);
                    cycle_length = 0;
                }
            }
        }
    }
    c->output[index] = encode_iteration(c->max_iter, c->bias);
}

/* render_alternate_reference_impl without planes, renderer.cpp:12064-12448,
 * over a grid or over caller-supplied dc offsets (points, width = count,
 * height = 1). */
static int render_alternate(float *output, int width, int height, fe_t zoom, const fo_reference *r, int max_iter,
                            const FractalRenderOptions *o, uint64_t *stats, int threads, struct timespec started,
                            const sc_t *points, const fe_t *point_radius, const fo_rect *crop) {
    const fo_rect area = crop ? *crop : (fo_rect){0, 0, width, height};
    const double factor = viewport_height_factor(o->coordinate_mode);
    const fe_t view_height = fe_muld(fe_div(fe_from_parts(1.0, 0), zoom), factor);
    const fe_t view_width = fe_muld(view_height, (double)width / (double)height);
    fe_t *xo = (fe_t *)malloc(sizeof(fe_t) * (size_t)width);
    fe_t *yo = (fe_t *)malloc(sizeof(fe_t) * (size_t)height);
    if (!xo || !yo) { free(xo); free(yo); return 1; }
    for (int px = 0; px < width; ++px)
        xo[px] = fe_muld(view_width, pixel_axis_offset(px, width, o->coordinate_mode) / (double)width);
    for (int py = 0; py < height; ++py)
        yo[py] = fe_muld(view_height, -pixel_axis_offset(py, height, o->coordinate_mode) / (double)height);
    alt_ctx c;
    memset(&c, 0, sizeof(c));
    c.r = r; c.output = output; c.width = width; c.max_iter = max_iter; c.mode = o->escape_radius_mode;
    c.strict = o->strict != 0; c.bias = o->output_bias; c.xo = xo; c.yo = yo; c.points = points;
    c.crop_x = area.x; c.crop_y = area.y; c.out_width = area.width;
    c.cycle_enabled = o->disable_cycle == 0 && (o->strict == 0 || o->strict_cycle != 0);
    /* renderer.cpp:12121-12128 */
    c.series_available = r->formula == 1 && r->series_enabled && r->series_iteration < max_iter
        && (!points || fe_compare(*point_radius, fe_sqrt(r->series_radius_squared)) <= 0);
    c.bla_enabled = o->disable_bla == 0 && r->alt_levels > 0;
    int mb = o->max_linear_bla_length;
    c.max_block = mb < 2 ? 2 : (mb > 4096 ? 4096 : mb);
    atomic_int exceeded;
    atomic_init(&exceeded, 0);
    c.budget_exceeded = &exceeded;
    if (o->time_budget_ms > 0) {
        c.budget_enabled = 1;
        clock_gettime(CLOCK_MONOTONIC, &c.deadline);
        long long ns = (long long)c.deadline.tv_nsec + (long long)o->time_budget_ms * 1000000LL;
        c.deadline.tv_sec += (time_t)(ns / 1000000000LL);
        c.deadline.tv_nsec = (long)(ns % 1000000000LL);
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (long long lp = 0; lp < (long long)area.width * area.height; ++lp)
        alt_pixel(&c, area.x + (int)(lp % area.width), area.y + (int)(lp / area.width));
    if (stats) {
        struct timespec ended;
        clock_gettime(CLOCK_MONOTONIC, &ended);
        memset(stats, 0, sizeof(uint64_t) * FO_STATS_FIELDS);
        stats[0] = (uint64_t)area.width * (uint64_t)area.height;
        stats[ST_RENDER_NS] = (uint64_t)((ended.tv_sec - started.tv_sec) * 1000000000LL
                                         + (ended.tv_nsec - started.tv_nsec));
    }
    free(xo); free(yo);
    return 0;
}
