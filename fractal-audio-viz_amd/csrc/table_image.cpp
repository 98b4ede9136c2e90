// table_image.cpp -- flat table images and per-render launch parameters
// (renderer.cpp:12689-12783 for the viewport and offsets).
#include "table_image.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <stdexcept>

namespace fav {

namespace {

template <typename Step>
void flatten(const std::vector<std::vector<Step>>& levels, std::vector<Step>& flat, LevelDir& dir) {
    if (levels.size() > static_cast<size_t>(kMaxLevels))
        throw std::runtime_error("BLA hierarchy deeper than the device level directory");
    flat.clear();
    dir = LevelDir{};
    dir.count = static_cast<int32_t>(levels.size());
    for (size_t l = 0; l < levels.size(); ++l) {
        dir.offset[l] = static_cast<int64_t>(flat.size());
        dir.size[l] = static_cast<int32_t>(levels[l].size());
        flat.insert(flat.end(), levels[l].begin(), levels[l].end());
    }
}

// Operand bounds of the plain-double block loop: every nonzero component of
// A, B (and of the loop's scaled delta and dc, checked on the device) lies in
// [2^-450, 2^500], and the smaller nonzero component of a value is at least
// 2^-420 times the larger.  Under these bounds every product, sum and
// cancellation of the block step A*delta + B*dc is a normal double both in
// plain form and in the reference's per-operand normalised form, so scaling
// by powers of two commutes with every rounding.
bool plain_operand(double re, double im, int e) {
    const double r = std::ldexp(re, e), i = std::ldexp(im, e);
    const double hi = std::max(std::fabs(r), std::fabs(i)), lo = std::min(std::fabs(r), std::fabs(i));
    if (!std::isfinite(hi) || hi > 0x1p500 || hi < 0x1p-450) return false;
    return lo == 0.0 || (lo >= 0x1p-450 && lo >= hi * 0x1p-420);
}

// x * 2^e as a plain double, 0 below the normal range and +inf above it (a
// validity radius, |Z|^2 or the negligible-delta bound of the block loop,
// compared with |delta|^2 values that the loop keeps in the normal range).
double plain_threshold(double m, int e) {
    if (!std::isfinite(m)) return HUGE_VAL;
    if (m == 0.0) return 0.0;
    const double v = std::ldexp(m, e);
    return std::fabs(v) < 0x1p-1022 ? 0.0 : v;
}

// Lookup and fast-path flags of flattened linear steps (LinearStepDev::pad[0]).
void mark_linear_steps(std::vector<LinearStepDev>& flat) {
    for (LinearStepDev& s : flat) {
        int32_t f = 0;
        if (s.length > 1 && std::isfinite(s.r2_m) && s.r2_m != 0.0) f |= kStepUsable;
        if (sc_normal(SC{s.a_re, s.a_im, s.a_e}) && sc_normal(SC{s.b_re, s.b_im, s.b_e})) f |= kStepNormalCoefficients;
        s.pad[0] = f;
    }
}

}  // namespace

void make_table_image(const ReferenceContext& c, TableImage& img) {
    const OrbitData& o = *c.orbit;
    const size_t n = o.scaled.size();
    img.re.resize(n);
    img.im.resize(n);
    img.e.resize(n);
    for (size_t i = 0; i < n; ++i) {
        img.re[i] = o.scaled[i].re;
        img.im[i] = o.scaled[i].im;
        img.e[i] = o.scaled[i].e;
    }
    img.fast.assign(n, OrbitFastDev{});
    if (o.zn_m.size() == n) {
        for (size_t i = 0; i < n; ++i) {
            OrbitFastDev& f = img.fast[i];
            f.z_re = o.scaled[i].re;
            f.z_im = o.scaled[i].im;
            f.z_e = o.scaled[i].e;
            f.zn_m = o.zn_m[i];
            f.zn_e = o.zn_e[i];
            f.margin_m = o.margin_m[i];
            f.margin_e = o.margin_e[i];
            f.small_exp = o.small_exp[i];
            f.flags = o.flags[i];
        }
    }
    flatten(c.cubic.levels, img.cubic, img.cubic_dir);
    for (int t = 0; t < 2; ++t) {
        flatten(c.linear[t].levels, img.linear[t], img.linear_dir[t]);
        mark_linear_steps(img.linear[t]);
    }
    img.series_re.clear();
    img.series_im.clear();
    img.series_e.clear();
    for (const SC& k : c.series_coefficients) {
        img.series_re.push_back(k.re);
        img.series_im.push_back(k.im);
        img.series_e.push_back(k.e);
    }
    flatten(c.alt.levels, img.alt, img.alt_dir);
}

// The high word of a double whose low word is 0 and whose value is at most
// v (v >= 0 or -0.0: the bits are cut), or at least v (v >= 0: rounded up
// through the carry into the high word, which stays a finite double for the
// bounds here).
static int32_t high_word_down(double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    return static_cast<int32_t>(static_cast<uint32_t>(bits >> 32));
}

static int32_t high_word_up(double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    const uint32_t hi = static_cast<uint32_t>(bits >> 32);
    return static_cast<int32_t>((bits & 0xffffffffULL) != 0 ? hi + 1 : hi);
}

// The rung table of one linear tier for the level cap top_cap
// (fav_device.h).
void make_plain_tables(const ReferenceContext& c, int tier, int top_cap, bool order3, std::vector<RungDev>& rungs) {
    const OrbitData& o = *c.orbit;
    const size_t n = o.scaled.size();
    rungs.clear();
    if (o.zn_m.size() != n || o.flags.size() != n || o.small_exp.size() != n) return;
    std::vector<LinearStepDev> flat;
    LevelDir dir{};
    flatten(c.linear[tier].levels, flat, dir);
    mark_linear_steps(flat);
    const int map_end = c.map_end;
    rungs.assign(n, RungDev{});
    // the product-gap bounds of each rung (0 and 0 where slot 0 does not
    // apply), for its own first block and the block that ends at it
    std::vector<double> gap_lo(n, 0.0), gap_hi(n, 0.0);
    for (size_t r = 0; r < n; ++r) {
        RungDev& g = rungs[r];
        // past the level-0 steps (lookup start > size[0], renderer.cpp:
        // 8816-8877) no block starts: stop there too
        const bool stop = (o.flags[r] & (kOrbitGlitchFlag | kOrbitEscapeFlag)) != 0
            || (map_end > 1 && r >= static_cast<size_t>(map_end)) || r == 0 || dir.count == 0
            || r > static_cast<size_t>(dir.size[0]);
        g.small_d = stop ? 0.0 : std::ldexp(1.0, o.small_exp[r]);
        g.zn_d = plain_threshold(o.zn_m[r], o.zn_e[r]);
        if (r == 0 || dir.count == 0) continue;
        const unsigned offset = static_cast<unsigned>(r - 1);
        int top = std::min(dir.count - 1, top_cap);
        if (offset != 0) top = std::min(top, __builtin_ctz(offset));
        g.top = top;
        for (int k = 0; k < kRungSlots; ++k) {
            const int level = top - k;
            if (level < 1) break;
            const int64_t index = offset >> level;
            if (index >= dir.size[level]) continue;  // slot stays 0: no lane fits
            const LinearStepDev& s = flat[static_cast<size_t>(dir.offset[level] + index)];
            if ((s.pad[0] & kStepUsable) == 0) continue;
            double r2 = plain_threshold(s.r2_m, s.r2_e);
            const size_t end = r + static_cast<size_t>(s.length);
            const bool plain = s.length == (1 << level) && plain_operand(s.a_re, s.a_im, s.a_e)
                && plain_operand(s.b_re, s.b_im, s.b_e) && end < n && (o.flags[end] & kOrbitEndpointFlag) == 0;
            if (plain) {
                const PlainStepDev v{std::ldexp(s.a_re, s.a_e), std::ldexp(s.a_im, s.a_e),
                                     std::ldexp(s.b_re, s.b_e), std::ldexp(s.b_im, s.b_e)};
                if (k == 0) {
                    g.a_re = v.a_re;
                    g.a_im = v.a_im;
                    g.b_re = v.b_re;
                    g.b_im = v.b_im;
                } else {
                    g.step_low[k - 1] = v;
                }
            } else {
                r2 = -r2;  // sign bit: fitting this step ends the plain loop
                if (r2 == 0.0) r2 = -0.0;
            }
            g.r2[k] = r2;
        }
        // the slot-0 loop's bounds (fav_device.h); 256 << top must fit the
        // 32-bit stride
        if (top >= 1 && top <= 23 && !std::signbit(g.r2[0]) && g.r2[0] > 0.0) {
            double fast = g.r2[0];
            if (g.zn_d < HUGE_VAL) fast = std::min(fast, std::nextafter(g.zn_d, HUGE_VAL));
            if (order3) fast = std::min(fast, 0x1p-161);
            // room for the slot-0 loop's |delta|^2, formed with one FMA
            g.fast_r2 = fast * (1.0 - 0x1p-48);
            g.stride = 256u << top;
            // |A m|^2 / |B c|^2 = |A|^2/|B|^2 * |m|^2/|c|^2 within 2^-116 ..
            // 2^116 (long double: no overflow); the bounds are kept inside
            // 2^-1000 .. 2^1000, which only narrows them
            const long double a2 = static_cast<long double>(g.a_re) * g.a_re + static_cast<long double>(g.a_im) * g.a_im;
            const long double b2 = static_cast<long double>(g.b_re) * g.b_re + static_cast<long double>(g.b_im) * g.b_im;
            if (a2 > 0 && b2 > 0) {
                const long double ratio = b2 / a2;
                gap_lo[r] = static_cast<double>(std::max(ratio * 0x1p-116L, 0x1p-600L));
                gap_hi[r] = static_cast<double>(std::min(ratio * 0x1p116L, 0x1p600L));
            }
        }
        g.gap_lo_own = high_word_up(gap_lo[r]);
        g.gap_hi_own = high_word_down(gap_hi[r]);
    }
    // the bounds at slot 0's end index, once every rung is set: the square
    // of its negligible bound, its fast bound and its product gap
    for (size_t r = 1; r < n; ++r) {
        const size_t end = r + (size_t(1) << rungs[r].top);
        const double small = end < n ? rungs[end].small_d : 0.0;
        const double bound = std::min(small * small * (1.0 - 0x1p-48), end < n ? rungs[end].fast_r2 : 0.0);
        rungs[r].end_hi = high_word_down(bound >= 0x1p-900 && bound <= 0x1p200 ? bound : -0.0);
        rungs[r].gap_lo_end = high_word_up(end < n ? gap_lo[end] : 0.0);
        rungs[r].gap_hi_end = high_word_down(end < n ? gap_hi[end] : 0.0);
    }
}

namespace {

// lookup_level with |delta|^2 = 0 (field_core.h lookup_level_zero) over a
// flattened table.
template <typename Step>
int lookup_zero(const std::vector<Step>& steps, const LevelDir& d, int start, int max_length) {
    if (start <= 0 || max_length <= 0 || d.count == 0 || start > d.size[0]) return -1;
    int h = 31 - __builtin_clz(static_cast<unsigned>(max_length));
    if (h > d.count - 1) h = d.count - 1;
    const int offset = start - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= d.size[level]) continue;
        const int64_t k = d.offset[level] + index;
        const Step& s = steps[static_cast<size_t>(k)];
        if (s.length > 1 && std::isfinite(s.r2_m) && s.r2_m != 0.0) return static_cast<int>(k);
    }
    return -1;
}

void set_walk(RefWalkDev& w, const LinearStepDev& s) {
    w.a_re = s.a_re; w.a_im = s.a_im; w.a_e = s.a_e;
    w.b_re = s.b_re; w.b_im = s.b_im; w.b_e = s.b_e;
}

void set_walk(RefWalkDev& w, const CubicStepDev& s) {
    w.a_re = s.re[0]; w.a_im = s.im[0]; w.a_e = s.e[0];
    w.b_re = s.re[1]; w.b_im = s.im[1]; w.b_e = s.e[1];
}

template <typename Step>
void fill_walk(const OrbitData& o, int map_end, const std::vector<Step>& steps, const LevelDir& d, int limit,
               std::vector<RefWalkDev>& walk) {
    const size_t n = o.flags.size();
    walk.assign(n, RefWalkDev{});
    for (size_t r = 0; r < n; ++r) {
        if ((o.flags[r] & (kOrbitGlitchFlag | kOrbitEscapeFlag)) != 0) continue;
        if (map_end > 1 && r >= static_cast<size_t>(map_end)) continue;
        const int k = lookup_zero(steps, d, static_cast<int>(r), limit);
        if (k < 0) continue;
        const Step& s = steps[static_cast<size_t>(k)];
        const size_t end = r + static_cast<size_t>(s.length);
        if (end >= n || (o.flags[end] & kOrbitEndpointFlag) != 0) continue;
        set_walk(walk[r], s);
        walk[r].length = s.length;
    }
}

}  // namespace

void make_ref_walk(const ReferenceContext& c, int table, int limit, std::vector<RefWalkDev>& walk) {
    walk.clear();
    const OrbitData& o = *c.orbit;
    if (o.flags.size() != o.scaled.size() || limit <= 0) return;
    if (table == 2) {
        std::vector<CubicStepDev> flat;
        LevelDir dir{};
        flatten(c.cubic.levels, flat, dir);
        fill_walk(o, c.map_end, flat, dir, limit, walk);
    } else {
        std::vector<LinearStepDev> flat;
        LevelDir dir{};
        flatten(c.linear[table].levels, flat, dir);
        fill_walk(o, c.map_end, flat, dir, limit, walk);
    }
}

void prepare_alternate(const ReferenceContext& c, const FieldRequest& r, PreparedAlternate& out) {
    const int w = r.width, h = r.height;
    const FractalRenderOptions& opt = r.options;
    const FE zoom = parse_zoom(r.zoom_text, c.precision_bits);
    const double factor = opt.coordinate_mode == 1 ? 4.0 : 2.8;
    const FE view_height = fe_muld(fe_div(fe_from_parts(1.0, 0), zoom), factor);
    const FE view_width = fe_muld(view_height, static_cast<double>(w) / static_cast<double>(h));
    out.offset_m.resize(static_cast<size_t>(w) + h);
    out.offset_e.resize(static_cast<size_t>(w) + h);
    const bool kalles = opt.coordinate_mode == 1;
    const double ox = kalles ? static_cast<double>(w / 2) : static_cast<double>(w - 1) / 2.0;
    const double oy = kalles ? static_cast<double>(h / 2) : static_cast<double>(h - 1) / 2.0;
    for (int px = 0; px < w; ++px) {
        const FE v = fe_muld(view_width, (static_cast<double>(px) - ox) / static_cast<double>(w));
        out.offset_m[static_cast<size_t>(px)] = v.m;
        out.offset_e[static_cast<size_t>(px)] = v.e;
    }
    for (int py = 0; py < h; ++py) {
        const FE v = fe_muld(view_height, -(static_cast<double>(py) - oy) / static_cast<double>(h));
        out.offset_m[static_cast<size_t>(w + py)] = v.m;
        out.offset_e[static_cast<size_t>(w + py)] = v.e;
    }
    AltRenderParams& p = out.params;
    p = AltRenderParams{};
    p.orbit_size = static_cast<int32_t>(c.orbit->scaled.size());
    p.formula = c.formula;
    p.parameter_re = c.parameter.re;
    p.parameter_im = c.parameter.im;
    p.parameter_e = c.parameter.e;
    p.alt_start = c.alt_start;
    p.input_r2_m = c.alt_input_radius_squared.m;
    p.input_r2_e = c.alt_input_radius_squared.e;
    p.series.count = static_cast<int32_t>(c.series_coefficients.size());
    p.series.iteration = c.series_iteration;
    p.series.radius2_m = c.series_radius_squared.m;
    p.series.radius2_e = c.series_radius_squared.e;
    // renderer.cpp:12121-12128: a point render also needs its farthest point
    // inside the series radius
    p.series_available = c.formula == 1 && c.series_enabled && c.series_iteration < r.max_iter
        && (!r.points || fe_compare(r.point_radius, fe_sqrt(c.series_radius_squared)) <= 0);
    p.width = w;
    p.height = h;
    p.max_iter = r.max_iter;
    p.escape_radius_mode = opt.escape_radius_mode;
    p.strict = opt.strict != 0;
    p.cycle_enabled = opt.disable_cycle == 0 && (opt.strict == 0 || opt.strict_cycle != 0);
    p.bla_enabled = opt.disable_bla == 0 && !c.alt.levels.empty();
    p.max_block = std::clamp(opt.max_linear_bla_length, 2, 4096);
    p.output_bias = opt.output_bias;
}

void prepare_render(const ReferenceContext& c, const FieldRequest& r, PreparedRender& out) {
    const int w = r.width, h = r.height;
    const FractalRenderOptions& opt = r.options;
    const FE zoom = parse_zoom(r.zoom_text, c.precision_bits);
    // viewport, renderer.cpp:12689-12705
    const double factor = c.coordinate_mode == 1 ? 4.0 : 2.8;
    const FE inverse_zoom = fe_div(fe_from_parts(1.0, 0), zoom);
    const FE view_height = fe_muld(inverse_zoom, factor);
    const FE view_width = fe_muld(view_height, static_cast<double>(w) / static_cast<double>(h));
    // a point render is one row of point_count values (renderer.cpp:20099-20101)
    const FE pixel_spacing = fe_div(view_height, fe_from_parts(r.points ? 1.0 : static_cast<double>(h), 0));
    // a point render uses the radius of its farthest point (renderer.cpp:12701-12705)
    const FE current_input_radius = r.points ? r.point_radius : fe_muld(inverse_zoom, factor);
    const bool covers = fe_compare(current_input_radius, c.input_radius) <= 0;
    const bool use_deep_linear = fe_compare(current_input_radius, c.deep_input_radius) <= 0;
    const bool use_ultra = fe_compare(current_input_radius, fe_muld(c.input_radius, 1.0e-40)) <= 0;
    // per-column / per-row offsets, renderer.cpp:12764-12783
    out.offset_m.resize(static_cast<size_t>(w) + h);
    out.offset_e.resize(static_cast<size_t>(w) + h);
    const bool kalles = opt.coordinate_mode == 1;
    const double ox = kalles ? static_cast<double>(w / 2) : static_cast<double>(w - 1) / 2.0;
    const double oy = kalles ? static_cast<double>(h / 2) : static_cast<double>(h - 1) / 2.0;
    for (int px = 0; px < w; ++px) {
        const FE v = fe_muld(view_width, (static_cast<double>(px) - ox) / static_cast<double>(w));
        out.offset_m[static_cast<size_t>(px)] = v.m;
        out.offset_e[static_cast<size_t>(px)] = v.e;
    }
    for (int py = 0; py < h; ++py) {
        const FE v = fe_muld(view_height, -(static_cast<double>(py) - oy) / static_cast<double>(h));
        out.offset_m[static_cast<size_t>(w + py)] = v.m;
        out.offset_e[static_cast<size_t>(w + py)] = v.e;
    }
    RenderParams& p = out.params;
    p = RenderParams{};
    p.orbit.size = static_cast<int32_t>(c.orbit->scaled.size());
    p.orbit.map_end = c.map_end;
    p.series.count = static_cast<int32_t>(c.series_coefficients.size());
    p.series.iteration = c.series_iteration;
    p.series.radius2_m = c.series_radius_squared.m;
    p.series.radius2_e = c.series_radius_squared.e;
    p.series.available = c.series_enabled && c.series_order >= opt.series_min_terms
        && c.series_order <= opt.series_max_terms
        && fe_compare(current_input_radius, fe_sqrt(c.series_radius_squared)) <= 0
        && c.series_iteration < r.max_iter;
    out.linear_tier = use_deep_linear ? 1 : 0;
    p.output_bias = opt.output_bias;
    p.pixel_spacing_m = pixel_spacing.m;
    p.pixel_spacing_e = pixel_spacing.e;
    p.width = w;
    p.height = h;
    p.max_iter = r.max_iter;
    p.escape_radius_mode = opt.escape_radius_mode;
    p.strict = opt.strict != 0;
    p.allow_recovery = opt.allow_recovery != 0;
    p.disable_bla = !covers || opt.disable_bla != 0;
    p.cycle_enabled = opt.disable_cycle == 0 && (opt.strict == 0 || opt.strict_cycle != 0);
    p.strict_cycle = opt.strict_cycle != 0;
    // renderer.cpp:12741-12750
    p.max_bla_length = std::clamp(std::min(r.series_block, opt.max_bla_length), 2, 64);
    p.max_linear_length = std::clamp(std::min(r.series_block, opt.max_linear_bla_length), 2,
                                     use_ultra ? 4096 : (use_deep_linear ? 1024 : 64));
    p.approximation_order = std::clamp(r.series_order, 1, 3);
    // the linear fast kernel skips the per-step budget clock and cycle checks
    p.fast_linear = !(p.cycle_enabled || opt.time_budget_ms > 0);
    p.reference_pixel = -1;
}

}  // namespace fav

namespace fav {

// Carves the per-pixel state arrays and queues out of one block of at least
// pixels * kStateBytesPerPixel + 4096 bytes.
void bind_state_store(RenderParams& p, void* block, size_t pixels) {
    char* cursor = static_cast<char*>(block);
    auto take = [&](size_t bytes) {
        char* here = cursor;
        cursor += (bytes + 255) & ~size_t(255);
        return here;
    };
    StateStore& q = p.state;
    double** doubles[] = {&q.d_re, &q.d_im, &q.g_re, &q.g_im, &q.t_re, &q.t_im, &q.tn_m, &q.pn_m, &q.mg_m};
    for (double** d : doubles) *d = reinterpret_cast<double*>(take(pixels * 8));
    int32_t** ints[] = {&q.d_e, &q.g_e, &q.t_e, &q.tn_e, &q.pn_e, &q.mg_e, &q.ref, &q.iter, &q.misc};
    for (int32_t** i : ints) *i = reinterpret_cast<int32_t*>(take(pixels * 4));
    q.general_queue = reinterpret_cast<uint32_t*>(take(pixels * 4));
    q.tail_queue = reinterpret_cast<uint32_t*>(take(pixels * 4));
    q.counters = reinterpret_cast<unsigned*>(take(64));
}

}  // namespace fav
