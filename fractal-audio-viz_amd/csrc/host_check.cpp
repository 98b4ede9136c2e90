// host_check.cpp -- TEST-ONLY CPU build of the field kernel's per-pixel code.
//
// Links the product's host table builders with field_core.h compiled for the
// CPU, so tests can check the kernel's logic (and any rewrite of it) against
// the oracle bit for bit without a GPU.  It is built into its own shared
// object (_build/libfav_host_check.so) that the product library never loads.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// Commits of the plain-double block loop, counted by the CPU build only
// (tests read how much of a field the loop covers).
static unsigned long long g_plain_blocks = 0;
#define FAV_PLAIN_BLOCK_PROBE() __atomic_fetch_add(&g_plain_blocks, 1ULL, __ATOMIC_RELAXED)
static unsigned long long g_plain_exits[16] = {};
#define FAV_PLAIN_EXIT_PROBE(reason) __atomic_fetch_add(&g_plain_exits[(reason) & 15], 1ULL, __ATOMIC_RELAXED)

#include "field_core.h"
#include "table_image.h"

// Tests switch the plain-double block loop off to compare both block loops
// against the same goldens.
static int g_plain_enabled = 1;
static bool plain_blocks_enabled() { return g_plain_enabled != 0; }

extern "C" void fav_host_check_set_plain(int enabled) { g_plain_enabled = enabled; }
// exits of the plain-double block loop by reason (field_core.h
// FAV_PLAIN_EXIT_PROBE codes); reset clears them
extern "C" unsigned long long fav_host_check_plain_exits(int reason, int reset) {
    const unsigned long long v = g_plain_exits[reason & 15];
    if (reset) g_plain_exits[reason & 15] = 0;
    return v;
}
extern "C" unsigned long long fav_host_check_plain_blocks(int reset) {
    const unsigned long long v = g_plain_blocks;
    if (reset) g_plain_blocks = 0;
    return v;
}

extern "C" int fav_host_check_render(
    const char* x, const char* y, const char* ref_zoom, int max_iter, int bits, int image_series_order,
    const char* clone_zoom, const FractalRenderOptions* options, float* output, int width, int height,
    const char* zoom, int series_order, int series_block, FractalRenderPlanes* planes, uint64_t* stats,
    char* error, int error_capacity) {
    using namespace fav;
    try {
        auto root = build_reference(x, y, ref_zoom, max_iter, bits, image_series_order, clone_zoom != nullptr, 0,
                                    options->escape_radius_mode, options->coordinate_mode);
        std::unique_ptr<ReferenceContext> clone;
        if (clone_zoom) clone = clone_reference(*root, clone_zoom);
        ReferenceContext& c = clone ? *clone : *root;
        TableImage img;
        make_table_image(c, img);
        FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, *options,
                             planes, stats, false};
        PreparedRender prep;
        prepare_render(c, request, prep);
        RenderParams p = prep.params;
        const OrbitData& o = *c.orbit;
        p.orbit.re = img.re.data();
        p.orbit.im = img.im.data();
        p.orbit.e = img.e.data();
        p.orbit.zr_m = o.zr_m.data();
        p.orbit.zi_m = o.zi_m.data();
        p.orbit.zr_e = o.zr_e.data();
        p.orbit.zi_e = o.zi_e.data();
        p.orbit.margin_m = o.margin_m.data();
        p.orbit.margin_e = o.margin_e.data();
        p.orbit.glitch_m = o.glitch_m.data();
        p.orbit.glitch_e = o.glitch_e.data();
        p.orbit.real_double = o.real_double.data();
        p.orbit.imag_double = o.imag_double.data();
        p.orbit.zn_m = o.zn_m.data();
        p.orbit.zn_e = o.zn_e.data();
        p.orbit.small_exp = o.small_exp.data();
        p.orbit.flags = o.flags.data();
        p.orbit.fast = img.fast.data();
        p.series.re = img.series_re.data();
        p.series.im = img.series_im.data();
        p.series.e = img.series_e.data();
        p.linear = img.linear[prep.linear_tier].data();
        p.linear_dir = img.linear_dir[prep.linear_tier];
        std::vector<RungDev> rungs;
        if (plain_blocks_enabled()) {
            p.plain_top = 31 - __builtin_clz(static_cast<unsigned>(p.max_linear_length));
            make_plain_tables(c, prep.linear_tier, p.plain_top, p.approximation_order >= 3, rungs);
            if (!rungs.empty()) p.rungs = rungs.data();
        }
        p.cubic = img.cubic.data();
        p.cubic_dir = img.cubic_dir;
        std::vector<RefWalkDev> walk;
        {
            const bool linear = p.approximation_order <= 1;
            const int level = 31 - __builtin_clz(static_cast<unsigned>(linear ? p.max_linear_length : p.max_bla_length));
            make_ref_walk(c, linear ? prep.linear_tier : 2, 1 << level, walk);
            if (!walk.empty()) {
                p.ref_walk = walk.data();
                p.ref_walk_limit = 1 << level;
            }
        }
        p.x_off_m = prep.offset_m.data();
        p.y_off_m = prep.offset_m.data() + width;
        p.x_off_e = prep.offset_e.data();
        p.y_off_e = prep.offset_e.data() + width;
        p.output = output;
        if (planes) {
            p.plane_iteration = planes->orbit_iteration;
            p.plane_phase = planes->phase;
            p.plane_de_x = planes->de_x;
            p.plane_de_y = planes->de_y;
            p.plane_test1 = planes->test1;
            p.plane_test2 = planes->test2;
        }
        std::vector<unsigned long long> counters(kStatsFields, 0);
        p.stats = counters.data();
        int flag = 0;
        unsigned long long start = 0;
        p.budget_flag = &flag;
        p.budget_start = &start;
        p.budget_ticks = 0;
        const long long n = static_cast<long long>(width) * height;
        std::vector<unsigned char> state_block(static_cast<size_t>(n) * kStateBytesPerPixel + 20 * 256);
        bind_state_store(p, state_block.data(), static_cast<size_t>(n));
        // The kernels of render_kernel.hip, run one after another on the
        // CPU: the plain pass (when the tables exist), the fast pass over
        // every pixel it did not finish, then the general queue, then the
        // tail queue.
        p.plain_pass = p.rungs != nullptr && p.fast_linear ? 1 : 0;
        std::vector<char> outcome(static_cast<size_t>(n), 0);
        auto pass = [&](auto&& body) {
#pragma omp parallel for schedule(dynamic, 64)
            for (long long i = 0; i < n; ++i) body(static_cast<size_t>(i));
        };
        auto run = [&](auto planes_tag, auto stats_tag) {
            constexpr bool Planes = decltype(planes_tag)::value;
            constexpr bool Stats = decltype(stats_tag)::value;
            if (p.plain_pass) {
                pass([&](size_t i) {
                    const int py = static_cast<int>(i / static_cast<size_t>(width));
                    const int px = static_cast<int>(i % static_cast<size_t>(width));
                    core::Counters<Stats> st;
                    st.clear();
                    const SC dc = core::pixel_dc(p, px, py);
                    core::PixelState s = core::init_pixel<Planes, Stats>(p, i, dc, st);
                    int32_t dlog[core::kDerivativeLog];
                    if (core::plain_pixel<Planes, Stats>(p, i, s, dc, st, dlog)) p.state.iter[i] = p.max_iter;
                    else core::save_state(p.state, i, s);
                    if constexpr (Stats) st.flush(p.stats);
                });
            }
            pass([&](size_t i) {
                if (p.plain_pass && p.state.iter[i] >= p.max_iter) return;
                const int py = static_cast<int>(i / static_cast<size_t>(width));
                const int px = static_cast<int>(i % static_cast<size_t>(width));
                core::Counters<Stats> st;
                st.clear();
                const SC dc = core::pixel_dc(p, px, py);
                core::PixelState s = p.plain_pass ? core::load_state(p.state, i)
                                                  : core::init_pixel<Planes, Stats>(p, i, dc, st);
                int32_t dlog[core::kDerivativeLog];
                if (core::run_linear<Planes, Stats>(p, i, s, dc, st, dlog) == core::kHandoff) {
                    core::save_state(p.state, i, s);
                    outcome[i] = 1;
                }
                if constexpr (Stats) st.flush(p.stats);
            });
            pass([&](size_t i) {
                if (outcome[i] != 1) return;
                const int py = static_cast<int>(i / static_cast<size_t>(width));
                const int px = static_cast<int>(i % static_cast<size_t>(width));
                core::Counters<Stats> st;
                st.clear();
                const SC dc = core::pixel_dc(p, px, py);
                core::PixelState s = core::load_state(p.state, i);
                if (core::run_general<Planes, Stats, false>(p, i, s, dc, st) == core::kTail) {
                    core::save_state(p.state, i, s);
                    outcome[i] = 2;
                }
                if constexpr (Stats) st.flush(p.stats);
            });
            pass([&](size_t i) {
                if (outcome[i] != 2) return;
                const int py = static_cast<int>(i / static_cast<size_t>(width));
                const int px = static_cast<int>(i % static_cast<size_t>(width));
                core::Counters<Stats> st;
                st.clear();
                const SC dc = core::pixel_dc(p, px, py);
                core::PixelState s = core::load_state(p.state, i);
                core::run_tail<Planes, Stats>(p, i, s, dc, st);
                if constexpr (Stats) st.flush(p.stats);
            });
        };
        if (planes && stats) run(std::true_type{}, std::true_type{});
        else if (planes) run(std::true_type{}, std::false_type{});
        else if (stats) run(std::false_type{}, std::true_type{});
        else run(std::false_type{}, std::false_type{});
        if (stats) for (int i = 0; i < kStatsFields; ++i) stats[i] = counters[static_cast<size_t>(i)];
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(error, static_cast<size_t>(error_capacity), "%s", e.what());
        return 1;
    }
}

#include "kfp_prepare.h"

// KFP colour pass of kfp_kernel.hip run on the CPU (test-only): the same
// option validation and per-pixel code, with the method-4 range reduced
// serially.  Returns 1 with `err` set when the options are rejected.
extern "C" int fav_host_check_kfp_planes(const float* field, uint8_t* output, int width, int height, int max_iter,
                                         const FractalKfpOptions* options, const FractalKfpPlanes* planes,
                                         const uint8_t* lut, int lut_size, char* err, int cap) {
    try {
        if (!fav::kfp::valid_options(options, lut_size) || !fav::kfp::valid_planes(planes, width, height))
            throw std::runtime_error("invalid KFP options");
        fav::kfp::KfpParams k =
            fav::kfp::make_params(*options, field, output, width, height, max_iter, lut, lut_size);
        fav::kfp::set_plane_context(k, *options, planes);
        if (planes) {
            k.orbit_iteration = planes->orbit_iteration;
            k.iteration = planes->iteration;
            k.bailout = planes->bailout;
            k.transition = planes->transition;
            k.phase = planes->phase;
            k.de_x = planes->de_x;
            k.de_y = planes->de_y;
            k.test1 = planes->test1;
            k.test2 = planes->test2;
            k.texture = k.texture_active ? planes->texture_rgb : nullptr;
        }
        double bounds[2] = {0.0, 1.0};
        if (k.color_method == 4) {
            long long lo = 0, hi = 0;
            bool any = false;
            for (size_t i = 0; i < static_cast<size_t>(width) * height; ++i) {
                long long v;
                if (!fav::kfp::transfer_bound_value(k, i, v)) continue;
                lo = any ? std::min(lo, v) : v;
                hi = any ? std::max(hi, v) : v;
                any = true;
            }
            if (any) {
                bounds[0] = static_cast<double>(lo);
                bounds[1] = static_cast<double>(hi);
            }
        }
        k.bounds = bounds;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) fav::kfp::colour_pixel(k, x, y);
        return 0;
    } catch (const std::exception& e) {
        if (err && cap > 0) {
            std::strncpy(err, e.what(), static_cast<size_t>(cap) - 1);
            err[cap - 1] = 0;
        }
        return 1;
    }
}

extern "C" int fav_host_check_kfp(const float* field, uint8_t* output, int width, int height, int max_iter,
                                  const FractalKfpOptions* options, const uint8_t* lut, int lut_size, char* err,
                                  int cap) {
    return fav_host_check_kfp_planes(field, output, width, height, max_iter, options, nullptr, lut, lut_size, err,
                                     cap);
}

#include "direct_prepare.h"

// Shallow direct renderer of direct_kernel.hip run on the CPU (test-only):
// the same set-up and per-pixel code.  Returns 1 with `err` set on a
// rejected argument.
extern "C" int fav_host_check_direct(float* output, int width, int height, const char* zoom, const char* x,
                                     const char* y, int max_iter, int formula, double julia_re, double julia_im,
                                     double output_bias, int escape_radius_mode, int coordinate_mode,
                                     uint32_t plane_flags, FractalRenderPlanes* planes, char* err, int cap) {
    try {
        fav::direct::DirectSetup d =
            fav::direct::make_setup(width, height, zoom, x, y, max_iter, formula, julia_re, julia_im, output_bias,
                                    escape_radius_mode, coordinate_mode, plane_flags);
        fav::direct::DirectParams p = d.params;
        p.x = d.x.data();
        p.y = d.y.data();
        p.output = output;
        if (planes) {
            p.plane_iteration = planes->orbit_iteration;
            p.plane_phase = planes->phase;
            p.plane_de_x = planes->de_x;
            p.plane_de_y = planes->de_y;
            p.plane_test1 = planes->test1;
            p.plane_test2 = planes->test2;
        }
        const bool de = planes && (plane_flags & fav::direct::kHintSkipAnalyticDe) == 0U;
#pragma omp parallel for schedule(dynamic, 1)
        for (int py = 0; py < height; ++py) {
            for (int px = 0; px < width; ++px) {
                switch (formula) {
                    case 0:
                        if (!planes) fav::direct::pixel<0, false, false>(p, px, py);
                        else if (de) fav::direct::pixel<0, true, true>(p, px, py);
                        else fav::direct::pixel<0, true, false>(p, px, py);
                        break;
                    case 1:
                        if (!planes) fav::direct::pixel<1, false, false>(p, px, py);
                        else if (de) fav::direct::pixel<1, true, true>(p, px, py);
                        else fav::direct::pixel<1, true, false>(p, px, py);
                        break;
                    case 2:
                        if (!planes) fav::direct::pixel<2, false, false>(p, px, py);
                        else fav::direct::pixel<2, true, false>(p, px, py);
                        break;
                    default:
                        if (!planes) fav::direct::pixel<3, false, false>(p, px, py);
                        else fav::direct::pixel<3, true, false>(p, px, py);
                        break;
                }
            }
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && cap > 0) {
            std::strncpy(err, e.what(), static_cast<size_t>(cap) - 1);
            err[cap - 1] = 0;
        }
        return 1;
    }
}

#include "alt_core.h"

// Alternate-formula field of alt_kernel.hip run on the CPU (test-only): the
// same reference build, tables and per-pixel code.
extern "C" int fav_host_check_alt(const char* x, const char* y, const char* ref_zoom, int max_iter, int bits,
                                  int formula, const char* julia_re, const char* julia_im, const char* clone_zoom,
                                  const FractalRenderOptions* options, float* output, int width, int height,
                                  const char* zoom, char* error, int error_capacity) {
    using namespace fav;
    try {
        auto root = build_reference(x, y, ref_zoom, max_iter, bits, 8, true, formula, options->escape_radius_mode,
                                    options->coordinate_mode, julia_re, julia_im);
        std::unique_ptr<ReferenceContext> clone;
        if (clone_zoom) clone = clone_reference(*root, clone_zoom);
        ReferenceContext& c = clone ? *clone : *root;
        TableImage img;
        make_table_image(c, img);
        FieldRequest request{output, width, height, zoom, max_iter, 8, 32, *options, nullptr, nullptr, false};
        PreparedAlternate prep;
        prepare_alternate(c, request, prep);
        AltRenderParams p = prep.params;
        const OrbitData& o = *c.orbit;
        p.re = img.re.data();
        p.im = img.im.data();
        p.e = img.e.data();
        p.margin_m = o.alt_margin_m.data();
        p.margin_e = o.alt_margin_e.data();
        p.steps = img.alt.data();
        p.dir = img.alt_dir;
        p.series.re = img.series_re.data();
        p.series.im = img.series_im.data();
        p.series.e = img.series_e.data();
        p.x_off_m = prep.offset_m.data();
        p.y_off_m = prep.offset_m.data() + width;
        p.x_off_e = prep.offset_e.data();
        p.y_off_e = prep.offset_e.data() + width;
        p.output = output;
        int flag = 0;
        unsigned long long start = 0;
        p.budget_flag = &flag;
        p.budget_start = &start;
        p.budget_ticks = 0;
#pragma omp parallel for schedule(dynamic, 16)
        for (int py = 0; py < height; ++py)
            for (int px = 0; px < width; ++px) alt::pixel(p, px, py);
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(error, static_cast<size_t>(error_capacity), "%s", e.what());
        return 1;
    }
}

#include <random>

namespace {

// A random normal ScaledComplex: larger part in [0.5, 1), the other part
// anywhere from exact zero through tiny to comparable, either sign.
fav::SC random_normal(std::mt19937_64& rng, int exponent_span) {
    std::uniform_real_distribution<double> unit(0.5, 1.0);
    std::uniform_int_distribution<int> pick(0, 7), expo(-exponent_span, exponent_span), small(-1100, 0);
    const double big = unit(rng) * (pick(rng) & 1 ? -1.0 : 1.0);
    double other = 0.0;
    switch (pick(rng)) {
        case 0: other = 0.0; break;
        case 1: other = -0.0; break;
        case 2: other = std::ldexp(unit(rng), small(rng)); break;
        case 3: other = big; break;
        case 4: other = -big; break;
        default: other = unit(rng) * 2.0 - 1.0; break;
    }
    const bool swap = pick(rng) & 1;
    return fav::SC{swap ? other : big, swap ? big : other, expo(rng)};
}

bool same_sc(fav::SC a, fav::SC b) {
    return std::memcmp(&a.re, &b.re, 8) == 0 && std::memcmp(&a.im, &b.im, 8) == 0 && a.e == b.e;
}

}  // namespace

// Test-only: the *_nn operations against the general ones on `count` random
// normal operand pairs (including exact cancellations and exponent gaps
// around the 60-binade cut).  Writes mismatch counts for mul, add, norm and
// the block loop's norm comparison (sn_at_least against sc_compare_norm on
// a normal norm and another norm that may be zero, infinite or equal).
extern "C" void fav_host_check_nn_ops(long long count, unsigned long long seed, long long* mismatches) {
    using namespace fav;
    std::mt19937_64 rng(seed);
    mismatches[0] = mismatches[1] = mismatches[2] = mismatches[3] = 0;
    std::uniform_int_distribution<int> gap(-70, 70), pick(0, 9);
    for (long long i = 0; i < count; ++i) {
        const SC a = random_normal(rng, 1 << 20);
        SC b = random_normal(rng, 1 << 20);
        if (pick(rng) < 6) b.e = a.e + gap(rng);  // exercise the alignment and the >60 cut
        if (pick(rng) == 0) b = SC{-a.re, -a.im, a.e};  // exact cancellation
        if (!sc_normal(a) || !sc_normal(b)) continue;
        if (!same_sc(sc_mul_nn(a, b), sc_mul(a, b))) ++mismatches[0];
        if (!same_sc(sc_add_nn(a, b), sc_add(a, b))) ++mismatches[1];
        const SN x = sc_norm_nn(a), y = sc_norm_squared(a);
        if (std::memcmp(&x.m, &y.m, 8) != 0 || x.e != y.e) ++mismatches[2];
        SN z = sc_norm_squared(b);
        const int kind = pick(rng);
        if (kind == 0) z = SN{0.0, 0};
        else if (kind == 1) z = SN{__builtin_inf(), INT_MAX};
        else if (kind == 2) z = x;
        else if (kind == 3) z.e = x.e;
        if (sn_at_least(z, x) != (sc_compare_norm(z, x) >= 0)) ++mismatches[3];
    }
}

#include "fx_core.h"

// The floatexp path's per-pixel code on the CPU (test-only), on the tables
// of a Mandelbrot reference built exactly as for the FP64 path.
extern "C" int fav_host_check_fx(const char* x, const char* y, const char* ref_zoom, int max_iter, int bits,
                                 const FractalRenderOptions* options, float* field, int32_t* counts, int width,
                                 int height, const char* zoom, int series_order, int series_block, char* error,
                                 int error_capacity) {
    using namespace fav;
    try {
        auto c = build_reference(x, y, ref_zoom, max_iter, bits, 8, false, 0, options->escape_radius_mode,
                                 options->coordinate_mode);
        TableImage img;
        make_table_image(*c, img);
        FieldRequest request{field, width, height, zoom, max_iter, series_order, series_block, *options, nullptr,
                             nullptr, false};
        PreparedRender prep;
        prepare_render(*c, request, prep);
        RenderParams p = prep.params;
        const OrbitData& o = *c->orbit;
        p.orbit.re = img.re.data();
        p.orbit.im = img.im.data();
        p.orbit.e = img.e.data();
        p.orbit.glitch_m = o.glitch_m.data();
        p.orbit.glitch_e = o.glitch_e.data();
        p.orbit.fast = img.fast.data();
        p.series.re = img.series_re.data();
        p.series.im = img.series_im.data();
        p.series.e = img.series_e.data();
        p.linear = img.linear[prep.linear_tier].data();
        p.linear_dir = img.linear_dir[prep.linear_tier];
        p.orbit.real_double = o.real_double.data();
        p.orbit.imag_double = o.imag_double.data();
        p.x_off_m = prep.offset_m.data();
        p.y_off_m = prep.offset_m.data() + width;
        p.x_off_e = prep.offset_e.data();
        p.y_off_e = prep.offset_e.data() + width;
        p.budget_ticks = 0;
        const fx::FxOut out{field, counts, 1};
#pragma omp parallel for schedule(dynamic, 16)
        for (int py = 0; py < height; ++py)
            for (int px = 0; px < width; ++px) fx::pixel(p, out, px, py);
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(error, static_cast<size_t>(error_capacity), "%s", e.what());
        return 1;
    }
}

// The explicit-rectangle KFP atlas of kfp_kernel.hip run on the CPU
// (test-only), with the caller's checks done by the test.
extern "C" int fav_host_check_kfp_rect(const float* parent, const float* child, int child_width, int child_height,
                                       uint8_t* output, int width, int height, int max_iter, int child_left,
                                       int child_top, int feather, const FractalKfpOptions* options,
                                       const uint8_t* lut, int lut_size) {
    try {
        fav::kfp::KfpParams k[2];
        k[0] = fav::kfp::make_params(*options, parent, output, width, height, max_iter, lut, lut_size);
        k[1] = child ? fav::kfp::make_params(*options, child, output, child_width, child_height, max_iter, lut,
                                             lut_size)
                     : fav::kfp::KfpParams{};
        double bounds[4] = {0.0, 1.0, 0.0, 1.0};
        for (int t = 0; t < 2; ++t) {
            k[t].bounds = bounds + 2 * t;
            if (k[t].color_method != 4 || !k[t].field) continue;
            long long lo = 0, hi = 0;
            bool any = false;
            for (size_t i = 0; i < static_cast<size_t>(k[t].width) * k[t].height; ++i) {
                long long v;
                if (!fav::kfp::transfer_bound_value(k[t], i, v)) continue;
                lo = any ? std::min(lo, v) : v;
                hi = any ? std::max(hi, v) : v;
                any = true;
            }
            if (any) {
                bounds[2 * t] = static_cast<double>(lo);
                bounds[2 * t + 1] = static_cast<double>(hi);
            }
        }
        std::vector<float> ex, ey;
        if (child && feather >= 2) {
            auto fill = [feather](std::vector<float>& v, int n) {
                v.resize(static_cast<size_t>(n));
                for (int i = 0; i < n; ++i) {
                    const int edge = std::min(i, n - 1 - i);
                    const float linear = std::min(1.0F, static_cast<float>(edge) / static_cast<float>(feather));
                    v[static_cast<size_t>(i)] = linear * linear * (3.0F - 2.0F * linear);
                }
            };
            fill(ex, child_width);
            fill(ey, child_height);
        }
        fav::kfp::RectAtlas g{child != nullptr, child_left, child_top, feather, ex.data(), ey.data()};
#pragma omp parallel for schedule(static)
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) fav::kfp::rect_atlas_pixel(k[0], k[1], g, x, y);
        return 0;
    } catch (...) {
        return 1;
    }
}

// The alternate-formula floatexp pixel (fx_core.h alt_pixel) on the CPU
// (test-only), on the tables of a reference built exactly as for the FP64
// alternate path, and the FP64 pixel (alt_core.h) of the same tables with
// its escape counts.
extern "C" int fav_host_check_fx_alt(const char* x, const char* y, const char* ref_zoom, int max_iter, int bits,
                                     int formula, const char* julia_re, const char* julia_im,
                                     const FractalRenderOptions* options, float* field, int32_t* counts,
                                     float* fp64_field, int32_t* fp64_counts, int width, int height,
                                     const char* zoom, char* error, int error_capacity) {
    using namespace fav;
    try {
        auto c = build_reference(x, y, ref_zoom, max_iter, bits, 8, true, formula, options->escape_radius_mode,
                                 options->coordinate_mode, julia_re, julia_im);
        TableImage img;
        make_table_image(*c, img);
        FieldRequest request{field, width, height, zoom, max_iter, 8, 32, *options, nullptr, nullptr, false};
        PreparedAlternate prep;
        prepare_alternate(*c, request, prep);
        AltRenderParams p = prep.params;
        const OrbitData& o = *c->orbit;
        p.re = img.re.data();
        p.im = img.im.data();
        p.e = img.e.data();
        p.margin_m = o.alt_margin_m.data();
        p.margin_e = o.alt_margin_e.data();
        p.steps = img.alt.data();
        p.dir = img.alt_dir;
        p.series.re = img.series_re.data();
        p.series.im = img.series_im.data();
        p.series.e = img.series_e.data();
        p.x_off_m = prep.offset_m.data();
        p.y_off_m = prep.offset_m.data() + width;
        p.x_off_e = prep.offset_e.data();
        p.y_off_e = prep.offset_e.data() + width;
        const fx::FxOut out{field, counts, 0};
        int flag = 0;
        unsigned long long start = 0;
        p.budget_flag = &flag;
        p.budget_start = &start;
        p.budget_ticks = 0;
        p.output = fp64_field;
        p.counts = fp64_counts;
#pragma omp parallel for schedule(dynamic, 16)
        for (int py = 0; py < height; ++py)
            for (int px = 0; px < width; ++px) {
                fx::alt_pixel(p, out, px, py);
                alt::pixel(p, px, py);
            }
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(error, static_cast<size_t>(error_capacity), "%s", e.what());
        return 1;
    }
}
