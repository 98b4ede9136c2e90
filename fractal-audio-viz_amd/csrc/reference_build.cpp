// reference_build.cpp -- host construction of a deep-zoom reference: the MPFR
// orbit, the image-wide parameter series and the cubic / linear BLA
// hierarchies (renderer.cpp:10640-10922, 11162-11868), written into the
// layouts the gfx950 kernel reads.
//
// Every table entry is produced with the same FloatExp operations, in the
// same order, as the reference builder, so the tables are bit-identical.
// Independent entries of one hierarchy level are built on several host
// threads; each entry is a pure function of the previous level, so the
// thread count does not change any bit.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <locale>
#include <sstream>
#include <stdexcept>
#include <thread>

#include <gmp.h>
#include <mpfr.h>

#include "fav_host.h"

namespace fav {
namespace {

constexpr int kMaxCubicLength = 64;          // MAX_SAFE_BLA_LENGTH, renderer.cpp:66
constexpr int kMaxLinearLength = 4096;       // MAX_SAFE_LINEAR_BLA_LENGTH, renderer.cpp:69

uint64_t now_ns() {
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now().time_since_epoch()).count());
}

template <typename Fn>
void parallel_for(size_t count, Fn fn) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>(std::min<size_t>(hw, 16), (count + 4095) / 4096);
    if (workers <= 1) {
        for (size_t i = 0; i < count; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (count + workers - 1) / workers;
    for (size_t w = 0; w < workers; ++w) {
        const size_t begin = w * chunk, end = std::min(count, begin + chunk);
        if (begin >= end) break;
        pool.emplace_back([=, &fn] { for (size_t i = begin; i < end; ++i) fn(i); });
    }
    for (auto& t : pool) t.join();
}

double viewport_height_factor(int coordinate_mode) {  // renderer.cpp:153-158
    return coordinate_mode == 1 ? 4.0 : 2.8;
}

// FloatExp::from_mpfr, renderer.cpp:7570-7579.
FE fe_from_mpfr(mpfr_srcptr value) {
    long exponent = 0;
    const double mantissa = mpfr_get_d_2exp(&exponent, value, MPFR_RNDN);
    if (!std::isfinite(mantissa)) return FE{mantissa, 0};
    if (exponent > INT_MAX) return FE{std::copysign(HUGE_VAL, mantissa), 0};
    if (exponent < INT_MIN) return FE{0.0, 0};
    return fe_from_parts(mantissa, static_cast<int>(exponent));
}

struct Mpfr {
    mpfr_t v;
    explicit Mpfr(long bits) { mpfr_init2(v, bits); }
    ~Mpfr() { mpfr_clear(v); }
    Mpfr(const Mpfr&) = delete;
    Mpfr& operator=(const Mpfr&) = delete;
    operator mpfr_ptr() { return v; }
    operator mpfr_srcptr() const { return v; }
};

// parse_coordinate, renderer.cpp:228-241 (locale-neutral long double).
long double parse_long(const char* text, const char* label) {
    std::istringstream stream(text ? text : "");
    stream.imbue(std::locale::classic());
    long double value = 0.0L;
    stream >> value;
    bool ok = !stream.fail();
    if (ok) {
        stream >> std::ws;
        ok = stream.eof();
    }
    if (!ok || !std::isfinite(value)) throw std::runtime_error(std::string("invalid ") + label + " coordinate");
    return value;
}

// Largest exponent t such that every normalised perturbation d with
// d.e <= t leaves the loop-top values of orbit index n unchanged:
// sc_add(Z, d) returns Z, and the escape margin
//   fe_add(fe_add(M, cross), dn),  cross = 2(zr*dr + zi*di),  dn = dr^2 + di^2
// (renderer.cpp:8383-8431) returns M = |Z|^2 - R^2.  Each FloatExp sum
// returns its larger operand unchanged when the exponents differ by more
// than 60 (renderer.cpp:8209, 7688), and for normalised operands
//   cross.e <= Z.e + d.e + 2   and   dn.e <= 2 d.e + 1
// (a product's exponent is at most the sum of its factors', a sum's at most
// one above the larger addend's, and the parts of Z and d have exponents at
// most Z.e and d.e).  So the three conditions
//   Z.e - d.e > 60,  M.e - (Z.e + d.e + 2) > 60,  M.e - (2 d.e + 1) > 60
// make the loop top take Z, M and R^2 + M exactly.  INT_MIN when Z or M is
// zero or the exponents are too large to bound safely.
int32_t negligible_delta_exponent(SC z, FE margin) {
    if ((z.re == 0.0 && z.im == 0.0) || margin.m == 0.0 || !fe_finite(margin) || !sc_finite(z)) return INT_MIN;
    const long long limit = 1LL << 28;
    if (z.e > limit || z.e < -limit || margin.e > limit || margin.e < -limit) return INT_MIN;
    const long long a = static_cast<long long>(z.e) - 61;
    const long long b = static_cast<long long>(margin.e) - z.e - 63;
    const long long c2 = static_cast<long long>(margin.e) - 62;  // 2t <= c2
    const long long c = c2 >= 0 ? c2 / 2 : -((-c2 + 1) / 2);     // floor(c2 / 2)
    long long t = a < b ? a : b;
    if (c < t) t = c;
    if (t < -limit) t = -limit;
    return static_cast<int32_t>(t);
}

// Repelling fixed point of z^2 + c, renderer.cpp:10736-10815: the root of
// z^2 - z + c with the larger modulus, used as the Julia reference start when
// the centre lies within 10^-digits of it (digits = max(32, ceil(log10 zoom)
// + 16)).  Returns true and sets zr, zi when it applies.
bool julia_fixed_point(const Mpfr& cx, const Mpfr& cy, const Mpfr& pr, const Mpfr& pi, const Mpfr& vz,
                       Mpfr& zr, Mpfr& zi, long bits) {
    Mpfr tmp(bits), dre(bits), dim(bits), norm(bits), dmag(bits), rre(bits), rim(bits), are(bits), aim(bits),
        bre(bits), bim(bits), anorm(bits), bnorm(bits), nr(bits), ni(bits), mult(bits), vlog(bits), tol(bits),
        dist(bits);
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
    if (mpfr_sgn(dim.v) < 0) mpfr_neg(rim, rim, MPFR_RNDN);
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
        mpfr_set(nr, are.v, MPFR_RNDN); mpfr_set(ni, aim.v, MPFR_RNDN); mpfr_set(norm, anorm.v, MPFR_RNDN);
    } else {
        mpfr_set(nr, bre.v, MPFR_RNDN); mpfr_set(ni, bim.v, MPFR_RNDN); mpfr_set(norm, bnorm.v, MPFR_RNDN);
    }
    mpfr_mul_ui(mult, norm, 4, MPFR_RNDN);
    mpfr_log10(vlog, vz, MPFR_RNDN);
    const double log_zoom = mpfr_get_d(vlog, MPFR_RNDN);
    const int digits = std::max(32, static_cast<int>(std::ceil(std::max(0.0, log_zoom))) + 16);
    mpfr_set_ui(tol, 10, MPFR_RNDN);
    mpfr_pow_si(tol, tol, -static_cast<long>(digits), MPFR_RNDN);
    mpfr_sub(tmp, cx, nr, MPFR_RNDN);
    mpfr_mul(dist, tmp, tmp, MPFR_RNDN);
    mpfr_sub(tmp, cy, ni, MPFR_RNDN);
    mpfr_mul(norm, tmp, tmp, MPFR_RNDN);
    mpfr_add(dist, dist, norm, MPFR_RNDN);
    mpfr_mul(norm, tol, tol, MPFR_RNDN);
    if (mpfr_cmp_ui(mult, 1) > 0 && mpfr_cmp(dist, norm) <= 0) {
        mpfr_set(zr, nr.v, MPFR_RNDN);
        mpfr_set(zi, ni.v, MPFR_RNDN);
        return true;
    }
    return false;
}

// make_reference_orbit, renderer.cpp:10640-10922: the MPFR orbit of the
// handle's formula (0 Mandelbrot, 1 Julia, 2 burning ship, 3 tricorn).
void make_orbit(ReferenceContext& c, const char* x, const char* y, const char* zoom, int max_iter,
                const char* julia_re, const char* julia_im, std::vector<FE>& alt_margins) {
    const long bits = c.precision_bits;
    const int formula = c.formula;
    Mpfr cx(bits), cy(bits), vz(bits), vr(bits), zr(bits), zi(bits), nr(bits), ni(bits), tmp(bits), pr(bits),
        pi(bits), ar(bits), ai(bits), norm(bits), margin(bits);
    if (mpfr_set_str(cx, x, 10, MPFR_RNDN) != 0 || mpfr_set_str(cy, y, 10, MPFR_RNDN) != 0)
        throw std::runtime_error("invalid MPFR Mandelbrot centre");
    if (zoom && mpfr_set_str(vz, zoom, 10, MPFR_RNDN) == 0 && mpfr_sgn(vz.v) > 0) {
        mpfr_set_d(vr, viewport_height_factor(c.coordinate_mode), MPFR_RNDN);
        mpfr_div(vr, vr, vz, MPFR_RNDN);
        c.input_radius = fe_from_mpfr(vr.v);
    } else {
        c.input_radius = fe_from_parts(1.0, 0);
    }
    if (formula == 1) {
        if (mpfr_set_str(pr, julia_re, 10, MPFR_RNDN) != 0 || mpfr_set_str(pi, julia_im, 10, MPFR_RNDN) != 0)
            throw std::runtime_error("invalid MPFR Julia constant");
        mpfr_set(zr, cx.v, MPFR_RNDN);
        mpfr_set(zi, cy.v, MPFR_RNDN);
    } else {
        mpfr_set(pr, cx.v, MPFR_RNDN);
        mpfr_set(pi, cy.v, MPFR_RNDN);
        mpfr_set_zero(zr, 0);
        mpfr_set_zero(zi, 0);
    }
    c.parameter = sc_from_fe(fe_from_mpfr(pr.v), fe_from_mpfr(pi.v));
    const bool fixed_point = formula == 1 && julia_fixed_point(cx, cy, pr, pi, vz, zr, zi, bits);
    c.fast_orbit.clear();
    c.fast_orbit.reserve(static_cast<size_t>(max_iter) + 1);
    alt_margins.clear();
    const double r2_double = escape_r2_double(c.escape_radius_mode);
    FE ore = fe_from_mpfr(zr.v), oim = fe_from_mpfr(zi.v);
    for (int i = 0; i <= max_iter; ++i) {
        if (!fe_finite(ore) || !fe_finite(oim)) break;
        c.fast_orbit.push_back(FEC{ore, oim});
        if (formula != 0) {  // |Z|^2 - R^2 in MPFR (only the alternate path reads it)
            mpfr_mul(norm, zr, zr, MPFR_RNDN);
            mpfr_mul(tmp, zi, zi, MPFR_RNDN);
            mpfr_add(norm, norm, tmp, MPFR_RNDN);
            mpfr_set_d(tmp, r2_double, MPFR_RNDN);
            mpfr_sub(margin, norm, tmp, MPFR_RNDN);
            alt_margins.push_back(fe_from_mpfr(margin.v));
        }
        if (fixed_point) continue;
        // each MPFR operation correctly rounded (RNDN)
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
        mpfr_swap(zr, nr);
        mpfr_swap(zi, ni);
        ore = fe_from_mpfr(zr.v);
        oim = fe_from_mpfr(zi.v);
    }
    const int n = static_cast<int>(c.fast_orbit.size());
    if (n == 0) throw std::runtime_error("reference orbit lost finite state at iteration zero");
    c.map_end = n - 1;
    const FE r2 = fe_escape_r2(c.escape_radius_mode);
    for (int i = 1; i < n; ++i) {
        if (fe_compare(fec_norm_squared(c.fast_orbit[static_cast<size_t>(i)]), r2) > 0) {
            c.map_end = i;
            break;
        }
    }
    auto orbit = std::make_shared<OrbitData>();
    orbit->scaled.resize(static_cast<size_t>(n));
    orbit->real_double.resize(static_cast<size_t>(n));
    orbit->imag_double.resize(static_cast<size_t>(n));
    orbit->zr_m.resize(static_cast<size_t>(n));
    orbit->zi_m.resize(static_cast<size_t>(n));
    orbit->margin_m.resize(static_cast<size_t>(n));
    orbit->glitch_m.resize(static_cast<size_t>(n));
    orbit->zr_e.resize(static_cast<size_t>(n));
    orbit->zi_e.resize(static_cast<size_t>(n));
    orbit->margin_e.resize(static_cast<size_t>(n));
    orbit->glitch_e.resize(static_cast<size_t>(n));
    orbit->zn_m.resize(static_cast<size_t>(n));
    orbit->zn_e.resize(static_cast<size_t>(n));
    orbit->small_exp.resize(static_cast<size_t>(n));
    orbit->flags.resize(static_cast<size_t>(n));
    const FE glitch_scale = fe_from_parts(1.0e-7, 0);
    const int mode = c.escape_radius_mode;
    parallel_for(static_cast<size_t>(n), [&](size_t i) {
        const SC z = sc_from_fe(c.fast_orbit[i].re, c.fast_orbit[i].im);
        orbit->scaled[i] = z;
        orbit->real_double[i] = ldexp_(z.re, z.e);
        orbit->imag_double[i] = ldexp_(z.im, z.e);
        // sc_component_as_float_exp of Z and (|Z|^2 - R^2), renderer.cpp:8401-8431
        const FEC parts = sc_parts(z);
        orbit->zr_m[i] = parts.re.m; orbit->zr_e[i] = parts.re.e;
        orbit->zi_m[i] = parts.im.m; orbit->zi_e[i] = parts.im.e;
        const FE margin = fe_sub(fec_norm_squared(parts), fe_escape_r2(mode));
        orbit->margin_m[i] = margin.m; orbit->margin_e[i] = margin.e;
        // glitch threshold 1e-7*|Z|^2 when |Z|^2 is nonzero and finite
        // (renderer.cpp:12931-12942); zero disables the test
        const SN zsq = sc_norm_squared(z);
        const FE reference_norm{zsq.m, zsq.e};
        FE threshold{0.0, 0};
        if (reference_norm.m != 0.0 && fe_finite(reference_norm))
            threshold = fe_mul(reference_norm, glitch_scale);
        orbit->glitch_m[i] = threshold.m; orbit->glitch_e[i] = threshold.e;
        const FE zn = fe_add(fe_escape_r2(mode), margin);  // sn_from_margin(margin)
        orbit->zn_m[i] = zn.m;
        orbit->zn_e[i] = zn.e;
        orbit->small_exp[i] = negligible_delta_exponent(z, margin);
        int32_t f = 0;
        if (threshold.m != 0.0 && fe_compare(zn, threshold) < 0) f |= kOrbitGlitchFlag;
        if (sc_outside_escape(SN{zn.m, zn.e}, mode) || fe_compare(margin, fe_zero()) > 0) f |= kOrbitEscapeFlag;
        if (sc_compare_norm(SN{zn.m, zn.e}, SN{0.75, 2}) >= 0) f |= kOrbitEndpointFlag;
        orbit->flags[i] = f;
    });
    if (formula != 0) {
        orbit->alt_margin_m.resize(static_cast<size_t>(n));
        orbit->alt_margin_e.resize(static_cast<size_t>(n));
        for (size_t i = 0; i < static_cast<size_t>(n); ++i) {
            orbit->alt_margin_m[i] = alt_margins[i].m;
            orbit->alt_margin_e[i] = alt_margins[i].e;
        }
        // renderer.cpp:10902-10920: repeat a settled cycle, then recompute
        // the map end from the repaired orbit
        stabilize_alternate_cycle(c, *orbit);
        c.map_end = n - 1;
        for (int i = 1; i < n; ++i) {
            if (fe_compare(fec_norm_squared(c.fast_orbit[static_cast<size_t>(i)]), r2) > 0) {
                c.map_end = i;
                break;
            }
        }
    }
    c.orbit = std::move(orbit);
}

// series_evaluate, renderer.cpp:11162-11174.
FEC series_evaluate(const std::vector<FEC>& k, FEC dc) {
    if (k.size() <= 1) return FEC{fe_zero(), fe_zero()};
    FEC result = k.back();
    for (size_t i = k.size() - 1; i > 1; --i) result = fec_add(fec_mul(result, dc), k[i - 1]);
    return fec_mul(result, dc);
}

// series_probe_is_safe, renderer.cpp:11176-11204.
bool series_probe_is_safe(FEC reference, FEC exact, FEC approximate, FE minimum_scale_squared,
                          FE tolerance_squared, int mode) {
    const FEC error = fec_sub(approximate, exact);
    FE scale = fec_norm_squared(exact);
    if (fe_compare(scale, minimum_scale_squared) < 0) scale = minimum_scale_squared;
    if (fe_compare(fec_norm_squared(error), fe_mul(scale, tolerance_squared)) > 0) return false;
    const bool exact_inside = fe_compare(fec_escape_margin_with_delta(reference, exact, mode), fe_zero()) <= 0;
    const bool approx_inside =
        fe_compare(fec_escape_margin_with_delta(reference, approximate, mode), fe_zero()) <= 0;
    return exact_inside && approx_inside;
}

// build_image_series, renderer.cpp:11206-11334 (Mandelbrot and Julia; a
// Julia pixel offset is the initial perturbation and c stays fixed).
void build_image_series(ReferenceContext& c, const std::vector<FEC>& orbit) {
    const bool julia = c.formula == 1;
    c.series_enabled = false;
    c.series_order = 0;
    c.series_iteration = 1;
    c.series_radius_squared = fe_zero();
    c.series_coefficients.clear();
    if (c.formula != 0 && !julia) return;
    if (orbit.size() < 16 || c.requested_max_iter < 8) return;
    const int order = std::clamp(c.requested_series_order, 8, 32);
    const int map_end = std::clamp(c.map_end, 2, static_cast<int>(orbit.size()) - 1);
    const int last_candidate = std::max(2, map_end - 1);
    const FE viewport_radius = c.input_radius;
    if (viewport_radius.m == 0.0 || !fe_finite(viewport_radius)) return;
    const FE probe_radius = fe_muld(viewport_radius, 1.5);
    const FE diagonal = fe_muld(probe_radius, 0.7071067811865476);
    const FE inner = fe_muld(probe_radius, 0.47);
    const FE z0 = fe_zero();
    const FEC probes[12] = {
        {probe_radius, z0}, {fe_neg(probe_radius), z0}, {z0, probe_radius}, {z0, fe_neg(probe_radius)},
        {diagonal, diagonal}, {fe_neg(diagonal), diagonal}, {diagonal, fe_neg(diagonal)},
        {fe_neg(diagonal), fe_neg(diagonal)}, {inner, fe_muld(inner, 0.63)},
        {fe_neg(inner), fe_muld(inner, 0.63)}, {fe_muld(inner, 0.63), inner},
        {fe_muld(inner, 0.63), fe_neg(inner)},
    };
    const size_t n = static_cast<size_t>(order) + 1;
    std::vector<FEC> coefficients(n, FEC{z0, z0}), next(n, FEC{z0, z0}), best;
    FEC exact[12];
    for (auto& e : exact) e = FEC{z0, z0};
    if (julia) {
        coefficients[1] = FEC{fe_from_parts(1.0, 0), z0};
        for (int p = 0; p < 12; ++p) exact[p] = probes[p];
    }
    int best_iteration = 1;
    const FE tolerance_squared = fe_from_parts(std::ldexp(1.0, -38), 0);
    const FE minimum_scale_squared = fe_sqr(probe_radius);
    const FE two = fe_from_parts(2.0, 0);
    const FEC one{fe_from_parts(1.0, 0), z0};
    // The reference's loop (renderer.cpp:11280-11322) advances the series
    // and then each probe's exact orbit by one iteration, and stops at the
    // first iteration where some probe is unsafe.  The series recursion
    // does not depend on the probes and each probe's orbit depends only on
    // itself, so the iterations are taken in chunks: while the twelve probes
    // check chunk i on their own threads, this thread computes the series of
    // chunk i + 1.  The stop is the earliest unsafe iteration of any probe,
    // and the kept series is the one after the iteration before it -- the
    // same iteration and coefficients as the sequential loop.  Chunks grow
    // from 64 to 2048 iterations, so a series that stops early computes
    // little beyond its stop.
    constexpr int kChunk = 2048;
    int chunk_size = 64;
    std::vector<FEC> chunk[2] = {std::vector<FEC>(static_cast<size_t>(kChunk) * n),
                                 std::vector<FEC>(static_cast<size_t>(kChunk) * n)};
    auto advance_series = [&](int first, int end, std::vector<FEC>& out) {
        for (int iteration = first; iteration < end; ++iteration) {
            const FEC reference = orbit[static_cast<size_t>(iteration)];
            std::fill(next.begin(), next.end(), FEC{z0, z0});
            for (int term = 1; term <= order; ++term) {
                FEC value = fec_mulf(fec_mul(reference, coefficients[static_cast<size_t>(term)]), two);
                for (int left = 1; left < term; ++left)
                    value = fec_add(value, fec_mul(coefficients[static_cast<size_t>(left)],
                                                   coefficients[static_cast<size_t>(term - left)]));
                if (!julia && term == 1) value = fec_add(value, one);
                next[static_cast<size_t>(term)] = value;
            }
            coefficients.swap(next);
            std::copy(coefficients.begin(), coefficients.end(),
                      out.begin() + static_cast<ptrdiff_t>(static_cast<size_t>(iteration - first) * n));
        }
    };
    int first = 0;
    int end = std::min(last_candidate, chunk_size);
    if (first < end) advance_series(first, end, chunk[0]);
    for (int index = 0; first < end; ++index) {
        const std::vector<FEC>& series_of = chunk[index & 1];
        int unsafe_at[12];
        auto run_probe = [&, first, end](int probe) {
            const FEC dc = probes[probe];
            std::vector<FEC> k(n);
            unsafe_at[probe] = end;
            for (int iteration = first; iteration < end; ++iteration) {
                const FEC reference = orbit[static_cast<size_t>(iteration)];
                const FEC exact_next = fec_add(fec_mulf(fec_mul(reference, exact[probe]), two),
                                               fec_add(fec_mul(exact[probe], exact[probe]),
                                                       julia ? FEC{z0, z0} : dc));
                exact[probe] = exact_next;
                if (iteration + 1 < order) continue;
                const FEC* series = series_of.data() + static_cast<size_t>(iteration - first) * n;
                k.assign(series, series + n);
                if (!series_probe_is_safe(orbit[static_cast<size_t>(iteration) + 1], exact_next,
                                          series_evaluate(k, dc), minimum_scale_squared, tolerance_squared,
                                          c.escape_radius_mode)) {
                    unsafe_at[probe] = iteration;
                    return;
                }
            }
        };
        // short chunks (the first few) are cheaper than starting threads
        const bool threaded = end - first >= 1024;
        std::vector<std::thread> pool;
        if (threaded) {
            for (int probe = 0; probe < 12; ++probe) pool.emplace_back(run_probe, probe);
        } else {
            for (int probe = 0; probe < 12; ++probe) run_probe(probe);
        }
        chunk_size = std::min(kChunk, chunk_size * 2);
        const int next_first = end;
        const int next_end = std::min(last_candidate, end + chunk_size);
        if (threaded && next_first < next_end) advance_series(next_first, next_end, chunk[(index + 1) & 1]);
        for (auto& t : pool) t.join();
        const int stop = *std::min_element(unsafe_at, unsafe_at + 12);
        // iterations first .. stop-1 were safe for every probe
        if (stop > first && stop >= order) {
            best_iteration = stop;
            const FEC* series = series_of.data() + static_cast<size_t>(stop - 1 - first) * n;
            best.assign(series, series + n);
        }
        if (stop < end) break;
        if (!threaded && next_first < next_end) advance_series(next_first, next_end, chunk[(index + 1) & 1]);
        first = next_first;
        end = next_end;
    }
    if (best_iteration <= 1 || best.empty()) return;
    c.series_enabled = true;
    c.series_order = order;
    c.series_iteration = best_iteration;
    c.series_radius_squared = fe_sqr(probe_radius);
    c.series_coefficients.reserve(best.size());
    for (const FEC& k : best) c.series_coefficients.push_back(sc_from_fe(k.re, k.im));
}

// The image series on its own thread, timed into series_build_ns, joined
// when this object ends.  The series and the BLA tables read the orbit and
// write disjoint fields of the context, so the caller builds the tables
// meanwhile.
class SeriesBuild {
public:
    SeriesBuild(ReferenceContext& c, const std::vector<FEC>& orbit)
        : thread_([&c, &orbit] {
              const uint64_t started = now_ns();
              build_image_series(c, orbit);
              c.series_build_ns = now_ns() - started;
          }) {}
    ~SeriesBuild() { thread_.join(); }
    SeriesBuild(const SeriesBuild&) = delete;
    SeriesBuild& operator=(const SeriesBuild&) = delete;

private:
    std::thread thread_;
};

// BlaStep (renderer.cpp:8476-8493) while the hierarchy is being composed.
struct CubicBuilder {
    FEC k[9];  // A B C D E F G H I
    FE r2;
    int length;
};

struct LinearBuilder {
    FEC a, b;
    FE r2;
    int length;
};

// Validity radius of the composition y∘x (renderer.cpp:11529-11539).
FE merged_radius(FEC x_a_coef, FEC x_b_coef, FE x_r2, FE y_r2, FE input_radius) {
    const FE x_a = fe_sqrt(fec_norm_squared(x_a_coef));
    const FE x_b = fe_sqrt(fec_norm_squared(x_b_coef));
    FE radius = fe_sqrt(x_r2);
    const FE remaining = fe_sub(fe_sqrt(y_r2), fe_mul(x_b, input_radius));
    if (fe_compare(remaining, fe_zero()) > 0 && fe_compare(x_a, fe_zero()) > 0)
        radius = fe_min(radius, fe_div(remaining, x_a));
    else
        radius = fe_zero();
    return radius;
}

// merge_bla, renderer.cpp:11528-11586: coefficients of y(x(d, c), c)
// truncated at total degree three.
CubicBuilder merge_cubic(const CubicBuilder& y, const CubicBuilder& x, FE input_radius) {
    const FE radius = merged_radius(x.k[0], x.k[1], x.r2, y.r2, input_radius);
    const FE two = fe_from_parts(2.0, 0), three = fe_from_parts(3.0, 0);
    const FEC &xA = x.k[0], &xB = x.k[1], &xC = x.k[2], &xD = x.k[3], &xE = x.k[4];
    const FEC &yA = y.k[0], &yB = y.k[1], &yC = y.k[2], &yD = y.k[3], &yE = y.k[4], &yF = y.k[5],
              &yG = y.k[6], &yH = y.k[7], &yI = y.k[8];
    const FEC x_a2 = fec_mul(xA, xA);
    const FEC x_ab = fec_mul(xA, xB);
    const FEC x_b2 = fec_mul(xB, xB);
    const FEC x_a3 = fec_mul(x_a2, xA);
    const FEC x_a2b = fec_mul(x_a2, xB);
    const FEC x_ab2 = fec_mul(xA, x_b2);
    const FEC x_b3 = fec_mul(x_b2, xB);
    CubicBuilder s;
    s.k[0] = fec_mul(yA, xA);
    s.k[1] = fec_add(fec_mul(yA, xB), yB);
    s.k[2] = fec_add(fec_mul(yA, xC), fec_mul(yC, x_a2));
    s.k[3] = fec_add(fec_add(fec_mul(yA, xD), fec_mul(yC, fec_mulf(x_ab, two))), fec_mul(yD, xA));
    s.k[4] = fec_add(fec_add(fec_mul(yA, xE), fec_mul(yC, x_b2)), fec_add(fec_mul(yD, xB), yE));
    s.k[5] = fec_add(fec_add(fec_mul(yA, x.k[5]), fec_mul(yC, fec_mulf(fec_mul(xA, xC), two))),
                     fec_mul(yF, x_a3));
    s.k[6] = fec_add(
        fec_add(fec_add(fec_mul(yA, x.k[6]),
                        fec_mul(yC, fec_mulf(fec_add(fec_mul(xA, xD), fec_mul(xB, xC)), two))),
                fec_add(fec_mul(yD, xC), fec_mul(yF, fec_mulf(x_a2b, three)))),
        fec_mul(yG, x_a2));
    s.k[7] = fec_add(
        fec_add(fec_add(fec_mul(yA, x.k[7]),
                        fec_mul(yC, fec_mulf(fec_add(fec_mul(xA, xE), fec_mul(xB, xD)), two))),
                fec_add(fec_mul(yD, xD), fec_mul(yF, fec_mulf(x_ab2, three)))),
        fec_add(fec_mul(yG, fec_mulf(x_ab, two)), fec_mul(yH, xA)));
    s.k[8] = fec_add(
        fec_add(fec_add(fec_mul(yA, x.k[8]), fec_mul(yC, fec_mulf(fec_mul(xB, xE), two))),
                fec_add(fec_mul(yD, xE), fec_mul(yF, x_b3))),
        fec_add(fec_mul(yG, x_b2), fec_add(fec_mul(yH, xB), yI)));
    s.r2 = fe_sqr(radius);
    s.length = x.length + y.length;
    return s;
}

// merge_linear_bla, renderer.cpp:11588-11610.
LinearBuilder merge_linear(const LinearBuilder& y, const LinearBuilder& x, FE input_radius) {
    const FE radius = merged_radius(x.a, x.b, x.r2, y.r2, input_radius);
    return LinearBuilder{fec_mul(y.a, x.a), fec_add(fec_mul(y.a, x.b), y.b), fe_sqr(radius),
                         x.length + y.length};
}

// compact_bla_step, renderer.cpp:8645-8671.
CubicStepDev compact_cubic(const CubicBuilder& b) {
    bool finite = fe_finite(b.r2);
    for (const FEC& k : b.k) finite = finite && fec_finite(k);
    CubicStepDev d{};
    for (int i = 0; i < 9; ++i) {
        const SC k = sc_from_fe(b.k[i].re, b.k[i].im);
        d.re[i] = k.re; d.im[i] = k.im; d.e[i] = k.e;
    }
    const FE r2 = finite ? b.r2 : fe_zero();
    d.r2_m = r2.m; d.r2_e = r2.e;
    d.length = b.length;
    return d;
}

// compact_linear_bla_step, renderer.cpp:8673-8684.
LinearStepDev compact_linear(const LinearBuilder& b) {
    const bool finite = fe_finite(b.r2) && fec_finite(b.a) && fec_finite(b.b);
    const SC a = sc_from_fe(b.a.re, b.a.im), bb = sc_from_fe(b.b.re, b.b.im);
    LinearStepDev d{};
    d.a_re = a.re; d.a_im = a.im; d.a_e = a.e;
    d.b_re = bb.re; d.b_im = bb.im; d.b_e = bb.e;
    const FE r2 = finite ? b.r2 : fe_zero();
    d.r2_m = r2.m; d.r2_e = r2.e;
    d.length = b.length;
    return d;
}

// build_linear_levels, renderer.cpp:11712-11757.
LinearLevels build_linear_levels(const ReferenceContext& c, int base_count, FE composition_radius) {
    const FE tolerance = fe_from_parts(std::ldexp(1.0, c.bla_epsilon_exponent), 0);
    const FE two = fe_from_parts(2.0, 0);
    std::vector<std::vector<LinearBuilder>> builder(1);
    builder[0].resize(static_cast<size_t>(base_count));
    parallel_for(static_cast<size_t>(base_count), [&](size_t i) {
        const FEC reference = c.fast_orbit[i + 1];
        const FE radius = fe_mul(tolerance, fe_sqrt(fec_norm_squared(reference)));
        builder[0][i] = LinearBuilder{fec_mulf(reference, two), FEC{fe_from_parts(1.0, 0), fe_zero()},
                                      fe_sqr(radius), 1};
    });
    for (size_t level = 1;; ++level) {
        if ((1ULL << level) > static_cast<unsigned long long>(kMaxLinearLength)) break;
        const size_t previous = builder[level - 1].size();
        const size_t current = (previous + 1) / 2;
        if (current == 0) break;
        builder.emplace_back(current);
        const auto& prev = builder[level - 1];
        auto& cur = builder[level];
        parallel_for(current, [&](size_t i) {
            const size_t first = i * 2;
            cur[i] = first + 1 < previous ? merge_linear(prev[first + 1], prev[first], composition_radius)
                                          : prev[first];
        });
    }
    LinearLevels out;
    out.levels.resize(builder.size());
    for (size_t l = 0; l < builder.size(); ++l) {
        out.levels[l].resize(builder[l].size());
        parallel_for(builder[l].size(), [&](size_t i) { out.levels[l][i] = compact_linear(builder[l][i]); });
    }
    return out;
}

// build_bla, Mandelbrot, renderer.cpp:11612-11769.
void build_bla(ReferenceContext& c) {
    const int max_iter = static_cast<int>(c.fast_orbit.size()) - 1;
    const int map_end = std::clamp(c.map_end, 0, max_iter);
    const int base_count = std::max(0, std::min(max_iter, map_end) - 1);
    const SeriesBuild series(c, c.fast_orbit);
    c.cubic.levels.clear();
    c.linear[0].levels.clear();
    c.linear[1].levels.clear();
    if (base_count == 0) return;
    const FE tolerance = fe_from_parts(std::ldexp(1.0, c.bla_epsilon_exponent), 0);
    const FE two = fe_from_parts(2.0, 0);
    const FEC zero{fe_zero(), fe_zero()};
    const FEC one{fe_from_parts(1.0, 0), fe_zero()};
    std::vector<std::vector<CubicBuilder>> builder(1);
    builder[0].resize(static_cast<size_t>(base_count));
    parallel_for(static_cast<size_t>(base_count), [&](size_t i) {
        const FEC reference = c.fast_orbit[i + 1];
        const FE radius = fe_mul(tolerance, fe_sqrt(fec_norm_squared(reference)));
        CubicBuilder s;
        s.k[0] = fec_mulf(reference, two);
        s.k[1] = one;
        s.k[2] = one;
        for (int k = 3; k < 9; ++k) s.k[k] = zero;
        s.r2 = fe_sqr(radius);
        s.length = 1;
        builder[0][i] = s;
    });
    for (size_t level = 1;; ++level) {
        if ((1ULL << level) > static_cast<unsigned long long>(kMaxCubicLength)) break;
        const size_t current = builder[level - 1].size() / 2;
        if (current == 0) break;
        builder.emplace_back(current);
        const auto& prev = builder[level - 1];
        auto& cur = builder[level];
        parallel_for(current, [&](size_t i) { cur[i] = merge_cubic(prev[i * 2 + 1], prev[i * 2], c.input_radius); });
    }
    c.cubic.levels.resize(builder.size());
    for (size_t l = 0; l < builder.size(); ++l) {
        c.cubic.levels[l].resize(builder[l].size());
        parallel_for(builder[l].size(), [&](size_t i) { c.cubic.levels[l][i] = compact_cubic(builder[l][i]); });
    }
    c.deep_input_radius = fe_muld(c.input_radius, 1.0e-20);
    c.linear[0] = build_linear_levels(c, base_count, c.input_radius);
    c.linear[1] = build_linear_levels(c, base_count, c.deep_input_radius);
}

// retarget_bla_radius, renderer.cpp:11776-11800.
FE retarget_radius(FE x_r2, FE y_r2, SC xa, SC xb, FE input_radius) {
    const SN na = sc_norm_squared(xa), nb = sc_norm_squared(xb);
    const FE x_a = fe_sqrt(FE{na.m, na.e});
    const FE x_b = fe_sqrt(FE{nb.m, nb.e});
    FE radius = fe_sqrt(x_r2);
    const FE remaining = fe_sub(fe_sqrt(y_r2), fe_mul(x_b, input_radius));
    if (fe_compare(remaining, fe_zero()) > 0 && fe_compare(x_a, fe_zero()) > 0)
        radius = fe_min(radius, fe_div(remaining, x_a));
    else
        radius = fe_zero();
    radius = fe_muld(radius, 1.0 - std::ldexp(1.0, -40));
    return fe_sqr(radius);
}

void retarget_linear(LinearLevels& t, FE input_radius) {  // renderer.cpp:11822-11845
    for (size_t level = 1; level < t.levels.size(); ++level) {
        const auto& prev = t.levels[level - 1];
        auto& cur = t.levels[level];
        parallel_for(cur.size(), [&](size_t i) {
            const size_t first = i * 2;
            if (first + 1 >= prev.size()) {
                cur[i].r2_m = prev[first].r2_m;
                cur[i].r2_e = prev[first].r2_e;
                return;
            }
            const LinearStepDev& x = prev[first];
            const LinearStepDev& y = prev[first + 1];
            const FE r2 = retarget_radius(FE{x.r2_m, x.r2_e}, FE{y.r2_m, y.r2_e}, SC{x.a_re, x.a_im, x.a_e},
                                          SC{x.b_re, x.b_im, x.b_e}, input_radius);
            cur[i].r2_m = r2.m;
            cur[i].r2_e = r2.e;
        });
    }
}

}  // namespace

FE parse_zoom(const char* text, long precision_bits) {  // renderer.cpp:11892-11915
    if (!text) throw std::runtime_error("native zoom text is too long or null");
    Mpfr value(precision_bits);
    const int status = mpfr_set_str(value, text, 10, MPFR_RNDN);
    if (status != 0 || mpfr_sgn(value.v) <= 0 || !mpfr_number_p(value.v))
        throw std::runtime_error("invalid deep Mandelbrot zoom");
    const FE result = fe_from_mpfr(value.v);
    if (!fe_finite(result) || result.m == 0.0)
        throw std::runtime_error("deep Mandelbrot zoom is outside the native exponent range");
    const long double log10_zoom =
        (std::log(std::fabs(static_cast<long double>(result.m)))
         + static_cast<long double>(result.e) * 0.693147180559945309417232121458176568L)
        / 2.302585092994045684017991454684364208L;
    if (!std::isfinite(log10_zoom) || log10_zoom < -300.0L || log10_zoom > 9800.0L)
        throw std::runtime_error("deep Mandelbrot zoom is outside the supported range");
    return result;
}

std::unique_ptr<ReferenceContext> build_reference(
    const char* x, const char* y, const char* zoom, int max_iter, int precision_bits, int series_order,
    bool reusable, int formula, int escape_radius_mode, int coordinate_mode, const char* julia_re,
    const char* julia_im, int bla_epsilon_exponent) {
    if (formula < 0 || formula > 3) throw std::runtime_error("invalid native reference formula or Julia constant");
    if (bla_epsilon_exponent < -60 || bla_epsilon_exponent > -20)
        throw std::runtime_error("BLA epsilon exponent must be in [-60, -20]");
    (void)parse_zoom(zoom, precision_bits);  // renderer.cpp:13638-13645
    auto c = std::make_unique<ReferenceContext>();
    c->formula = formula;
    c->escape_radius_mode = escape_radius_mode;
    c->bla_epsilon_exponent = bla_epsilon_exponent;
    c->coordinate_mode = coordinate_mode;
    c->precision_bits = std::max(128, precision_bits);
    c->requested_max_iter = max_iter;
    c->requested_series_order = std::clamp(series_order, 8, 32);
    c->reusable = reusable;
    c->x_text = x;
    c->y_text = y;
    c->x_center = parse_long(x, "real");
    c->y_center = parse_long(y, "imaginary");
    c->julia_re = static_cast<double>(parse_long(julia_re, "Julia real"));
    c->julia_im = static_cast<double>(parse_long(julia_im, "Julia imaginary"));
    const uint64_t t0 = now_ns();
    std::vector<FE> alt_margins;
    make_orbit(*c, x, y, zoom, max_iter, julia_re, julia_im, alt_margins);
    c->reference_build_ns = now_ns() - t0;
    const uint64_t t1 = now_ns();
    if (formula == 0) {
        build_bla(*c);
    } else {
        // renderer.cpp:13683-13697: Julia's image series, then the real
        // 2x2 hierarchy; alternate roots keep their builder orbit for clones
        {
            std::unique_ptr<SeriesBuild> series;
            if (formula == 1) series = std::make_unique<SeriesBuild>(*c, c->fast_orbit);
            build_alternate_bla(*c);
        }
        reusable = true;
        c->reusable = true;
    }
    c->bla_build_ns = now_ns() - t1;
    if (!reusable) {
        c->fast_orbit.clear();
        c->fast_orbit.shrink_to_fit();
    }
    return c;
}

// clone_reference_context + build_retargeted_bla, renderer.cpp:13704-13768,
// 11802-11868.
std::unique_ptr<ReferenceContext> clone_reference(const ReferenceContext& s, const char* zoom) {
    if (!s.orbit || s.fast_orbit.empty())
        throw std::runtime_error("reference was not created as a reusable tier root");
    const uint64_t t0 = now_ns();
    auto c = std::make_unique<ReferenceContext>();
    c->orbit = s.orbit;
    c->map_end = s.map_end;
    c->requested_max_iter = s.requested_max_iter;
    c->requested_series_order = s.requested_series_order;
    c->x_text = s.x_text;
    c->y_text = s.y_text;
    c->formula = s.formula;
    c->escape_radius_mode = s.escape_radius_mode;
    c->bla_epsilon_exponent = s.bla_epsilon_exponent;
    c->coordinate_mode = s.coordinate_mode;
    c->precision_bits = s.precision_bits;
    c->reference_build_ns = now_ns() - t0;
    const uint64_t t1 = now_ns();
    const FE parsed = parse_zoom(zoom, c->precision_bits);
    c->input_radius = fe_div(fe_from_parts(viewport_height_factor(c->coordinate_mode), 0), parsed);
    c->parameter = s.parameter;
    c->x_center = s.x_center;
    c->y_center = s.y_center;
    c->julia_re = s.julia_re;
    c->julia_im = s.julia_im;
    if (s.formula != 0) {  // renderer.cpp:13739-13760
        c->fast_orbit = s.fast_orbit;
        c->reusable = true;
        {
            std::unique_ptr<SeriesBuild> series;
            if (s.formula == 1) series = std::make_unique<SeriesBuild>(*c, c->fast_orbit);
            build_alternate_bla(*c);
        }
        c->bla_build_ns = now_ns() - t1;
        return c;
    }
    const SeriesBuild series(*c, s.fast_orbit);
    c->cubic = s.cubic;
    for (size_t level = 1; level < c->cubic.levels.size(); ++level) {
        const auto& prev = c->cubic.levels[level - 1];
        auto& cur = c->cubic.levels[level];
        parallel_for(cur.size(), [&](size_t i) {
            const CubicStepDev& x = prev[i * 2];
            const CubicStepDev& y = prev[i * 2 + 1];
            const FE r2 = retarget_radius(FE{x.r2_m, x.r2_e}, FE{y.r2_m, y.r2_e}, SC{x.re[0], x.im[0], x.e[0]},
                                          SC{x.re[1], x.im[1], x.e[1]}, c->input_radius);
            cur[i].r2_m = r2.m;
            cur[i].r2_e = r2.e;
        });
    }
    c->linear[0] = s.linear[0];
    retarget_linear(c->linear[0], c->input_radius);
    c->deep_input_radius = fe_muld(c->input_radius, 1.0e-20);
    c->linear[1] = s.linear[0];
    retarget_linear(c->linear[1], c->deep_input_radius);
    c->bla_build_ns = now_ns() - t1;
    return c;
}

}  // namespace fav
