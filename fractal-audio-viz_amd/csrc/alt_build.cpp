// alt_build.cpp -- host tables of the alternate deep formulas (Julia,
// burning ship, tricorn): the real 2x2 linear BLA hierarchy
// (build_alternate_linear_bla, renderer.cpp:11336-11526) and the repair of a
// settled projected cycle (stabilize_alternate_reference_cycle,
// renderer.cpp:9691-9788).  Every entry is formed with the reference's
// FloatExp operations in its order, so the tables are bit-identical.
#include <algorithm>
#include <cmath>

#include "fav_host.h"

namespace fav {
namespace {

struct AltBuilder {
    FE c[8];
    FE r2;
    int length;
};

FE fe_abs(FE v) { return FE{std::fabs(v.m), v.e}; }

FE matrix_norm(const FE* c, int offset) {  // alternate_matrix_norm, renderer.cpp:11340-11350
    FE sum{0.0, 0};
    for (int k = 0; k < 4; ++k) sum = fe_add(sum, fe_sqr(c[offset + k]));
    return fe_sqrt(sum);
}

// merge_alternate_linear_bla, renderer.cpp:11352-11397: M = My Mx,
// P = My Px + Py, radius bounded through the composed derivative.
AltBuilder merge(const AltBuilder& y, const AltBuilder& x, FE input_parameter_radius) {
    const FE* yc = y.c;
    const FE* xc = x.c;
    AltBuilder out{};
    out.c[0] = fe_add(fe_mul(yc[0], xc[0]), fe_mul(yc[1], xc[2]));
    out.c[1] = fe_add(fe_mul(yc[0], xc[1]), fe_mul(yc[1], xc[3]));
    out.c[2] = fe_add(fe_mul(yc[2], xc[0]), fe_mul(yc[3], xc[2]));
    out.c[3] = fe_add(fe_mul(yc[2], xc[1]), fe_mul(yc[3], xc[3]));
    out.c[4] = fe_add(fe_add(fe_mul(yc[0], xc[4]), fe_mul(yc[1], xc[6])), yc[4]);
    out.c[5] = fe_add(fe_add(fe_mul(yc[0], xc[5]), fe_mul(yc[1], xc[7])), yc[5]);
    out.c[6] = fe_add(fe_add(fe_mul(yc[2], xc[4]), fe_mul(yc[3], xc[6])), yc[6]);
    out.c[7] = fe_add(fe_add(fe_mul(yc[2], xc[5]), fe_mul(yc[3], xc[7])), yc[7]);
    const FE x_state = matrix_norm(xc, 0);
    const FE x_parameter = matrix_norm(xc, 4);
    FE radius = fe_sqrt(x.r2);
    const FE remaining = fe_sub(fe_sqrt(y.r2), fe_mul(x_parameter, input_parameter_radius));
    if (fe_compare(remaining, fe_zero()) > 0 && fe_compare(x_state, fe_zero()) > 0)
        radius = fe_min(radius, fe_div(remaining, x_state));
    else
        radius = fe_zero();
    radius = fe_muld(radius, 1.0 - std::ldexp(1.0, -40));
    out.r2 = fe_sqr(radius);
    out.length = x.length + y.length;
    return out;
}

AltStepDev compact(const AltBuilder& b) {
    AltStepDev s{};
    for (int k = 0; k < 8; ++k) {
        s.m[k] = b.c[k].m;
        s.e[k] = b.c[k].e;
    }
    s.r2_m = b.r2.m;
    s.r2_e = b.r2.e;
    s.length = b.length;
    return s;
}

}  // namespace

void build_alternate_bla(ReferenceContext& c) {
    c.alt.levels.clear();
    c.alt_input_radius_squared = fe_sqr(c.input_radius);
    c.alt_start = c.formula == 1 ? 0 : 1;
    const int orbit = static_cast<int>(c.fast_orbit.size());
    if (orbit < 2) return;
    const int first = c.alt_start;
    const int last = std::min(c.map_end, orbit - 1);
    const int base_count = last - first;
    if (base_count <= 0) return;
    const FE tolerance = fe_from_parts(std::ldexp(1.0, c.bla_epsilon_exponent), 0);
    const FE two = fe_from_parts(2.0, 0), quarter = fe_from_parts(0.25, 0), one = fe_from_parts(1.0, 0);
    std::vector<AltBuilder> base(static_cast<size_t>(base_count));
    for (int k = 0; k < base_count; ++k) {
        const FEC& ref = c.fast_orbit[static_cast<size_t>(first + k)];
        const FE re = ref.re, im = ref.im;
        AltBuilder b{};
        if (c.formula == 3) {  // conjugate map: d -> 2 conj(Z) conj(d)
            b.c[0] = fe_mul(re, two);
            b.c[1] = fe_muld(im, -2.0);
            b.c[2] = fe_muld(im, -2.0);
            b.c[3] = fe_muld(re, -2.0);
        } else if (c.formula == 2) {  // abs fold: signs of the reference's parts
            const FE are = fe_abs(re), aim = fe_abs(im);
            const double rs = re.m < 0.0 ? -1.0 : 1.0, is = im.m < 0.0 ? -1.0 : 1.0;
            b.c[0] = fe_muld(are, 2.0 * rs);
            b.c[1] = fe_muld(aim, -2.0 * is);
            b.c[2] = fe_muld(aim, 2.0 * rs);
            b.c[3] = fe_muld(are, 2.0 * is);
        } else {
            b.c[0] = fe_mul(re, two);
            b.c[1] = fe_muld(im, -2.0);
            b.c[2] = fe_mul(im, two);
            b.c[3] = fe_mul(re, two);
        }
        if (c.formula != 1) {
            b.c[4] = one;
            b.c[7] = one;
        }
        const FE magnitude = fe_sqrt(fec_norm_squared(ref));  // alternate_reference_radius
        FE radius = fe_mul(tolerance, fe_compare(magnitude, one) > 0 ? magnitude : one);
        if (c.formula == 2) {  // no multi-step map across an absolute-value axis
            if (re.m == 0.0 || im.m == 0.0) {
                radius = fe_zero();
            } else {
                radius = fe_min(radius, fe_mul(fe_abs(re), quarter));
                radius = fe_min(radius, fe_mul(fe_abs(im), quarter));
            }
        }
        b.r2 = fe_sqr(radius);
        b.length = 1;
        base[static_cast<size_t>(k)] = b;
    }
    std::vector<AltBuilder> previous = std::move(base);
    c.alt.levels.emplace_back();
    for (const AltBuilder& b : previous) c.alt.levels.back().push_back(compact(b));
    for (int level = 1; (1 << level) <= 4096; ++level) {  // MAX_SAFE_LINEAR_BLA_LENGTH
        const size_t size = (previous.size() + 1) / 2;
        if (size == 0) break;
        std::vector<AltBuilder> current(size);
        for (size_t i = 0; i < size; ++i) {
            const size_t f = i * 2;
            current[i] = f + 1 < previous.size() ? merge(previous[f + 1], previous[f], c.input_radius) : previous[f];
        }
        c.alt.levels.emplace_back();
        for (const AltBuilder& b : current) c.alt.levels.back().push_back(compact(b));
        previous = std::move(current);
    }
}

void stabilize_alternate_cycle(ReferenceContext& c, OrbitData& o) {
    const size_t count = std::min(c.fast_orbit.size(), o.scaled.size());
    if (count < 6 || o.real_double.size() != count || o.imag_double.size() != count) return;
    size_t last_finite = count;
    while (last_finite > 0) {
        const size_t index = last_finite - 1;
        if (std::isfinite(o.real_double[index]) && std::isfinite(o.imag_double[index])) break;
        --last_finite;
    }
    if (last_finite < 6) return;
    const size_t scan_limit = std::min<size_t>(last_finite - 1, 1024);
    const int max_period = std::min<int>(64, static_cast<int>((scan_limit + 1) / 4));
    const double tolerance = 1.0e-12;
    const double r2 = escape_r2_double(c.escape_radius_mode);
    for (int period = 1; period <= max_period; ++period) {
        for (size_t end = static_cast<size_t>(4 * period - 1); end <= scan_limit; ++end) {
            const size_t start = end - static_cast<size_t>(3 * period) + 1U;
            bool matches = true;
            for (int offset = 0; offset < period && matches; ++offset) {
                const size_t a = start + static_cast<size_t>(offset);
                const size_t b = a + static_cast<size_t>(period), d = b + static_cast<size_t>(period);
                const double ra = std::hypot(o.real_double[a], o.imag_double[a]);
                const double rb = std::hypot(o.real_double[b], o.imag_double[b]);
                const double rd = std::hypot(o.real_double[d], o.imag_double[d]);
                if (!std::isfinite(ra) || !std::isfinite(rb) || !std::isfinite(rd) || ra * ra >= r2
                    || rb * rb >= r2 || rd * rd >= r2) {
                    matches = false;
                    break;
                }
                const double scale = std::max(1.0, std::max(ra, std::max(rb, rd)));
                const double d1 = std::hypot(o.real_double[b] - o.real_double[a], o.imag_double[b] - o.imag_double[a]);
                const double d2 = std::hypot(o.real_double[d] - o.real_double[b], o.imag_double[d] - o.imag_double[b]);
                if (!(d1 <= tolerance * scale && d2 <= tolerance * scale)) matches = false;
            }
            if (!matches) continue;
            for (size_t target = start; target < count; ++target) {
                const size_t source = start + (target - start) % static_cast<size_t>(period);
                c.fast_orbit[target] = c.fast_orbit[source];
                o.scaled[target] = o.scaled[source];
                o.real_double[target] = o.real_double[source];
                o.imag_double[target] = o.imag_double[source];
                if (target < o.alt_margin_m.size() && source < o.alt_margin_m.size()) {
                    o.alt_margin_m[target] = o.alt_margin_m[source];
                    o.alt_margin_e[target] = o.alt_margin_e[source];
                }
            }
            return;
        }
    }
}

}  // namespace fav
