// field_core.h -- per-pixel deep-zoom state machine of the FP64 field path.
//
// Each pixel runs the reference's loop (renderer.cpp:12817-13544): image
// series jump, then glitch test, escape test, optional Brent cycle check,
// rebase, reference-escape switch, BLA lookup with endpoint replay,
// plain-double tail and exact perturbation steps.  Every floating-point
// operation on the control path is an IEEE add/sub/mul/compare or an exact
// frexp/ldexp, compiled with -ffp-contract=off, so escape counts are
// bit-identical to the scalar CPU path.
//
// The code is written once for the gfx950 kernel (render_kernel.hip) and
// compiles unchanged as host C++ for a test harness that checks the kernel
// logic against the oracle without a GPU (host_check.cpp, test-only).
#pragma once

#include "fav_device.h"
#include "fav_numeric.h"

#if !defined(__HIP_DEVICE_COMPILE__)
#include <atomic>
#include <chrono>
#endif

namespace fav {
namespace core {

constexpr double kLn2 = 0.693147180559945309417232121458176568;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;

enum StatField {  // RenderStats order, renderer.cpp:4615-4643
    ST_PIXELS, ST_LOGICAL, ST_BLA_BLOCKS, ST_LINEAR_BLOCKS, ST_CUBIC_BLOCKS, ST_EXACT,
    ST_REPLAY, ST_BLA_RETRIES, ST_CYCLE_INSIDE, ST_TAIL_PIXELS, ST_BLA_DISABLED,
    ST_TAIL_STEPS, ST_MAX_TAIL_STEPS, ST_TAIL_REBASES, ST_TAIL_FALLBACKS, ST_MAX_PIXEL_ITER,
    ST_SERIES_PIXELS, ST_SERIES_JUMPS, ST_GLITCH, ST_UNRESOLVED, ST_DEADLINE, ST_SECONDARY,
    ST_RENDER_NS, ST_HISTOGRAM
};

// ---- platform hooks ------------------------------------------------------
FAV_HD unsigned long long clock_ticks() {
#if defined(__HIP_DEVICE_COMPILE__)
    return wall_clock64();
#else
    return static_cast<unsigned long long>(std::chrono::duration_cast<std::chrono::microseconds>(
        std::chrono::steady_clock::now().time_since_epoch()).count()) * 100ULL;  // 100 MHz ticks
#endif
}

FAV_HD int flag_load(int* flag) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    return reinterpret_cast<std::atomic<int>*>(flag)->load(std::memory_order_relaxed);
#endif
}

FAV_HD void flag_store(int* flag) {
#if defined(__HIP_DEVICE_COMPILE__)
    __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    reinterpret_cast<std::atomic<int>*>(flag)->store(1, std::memory_order_relaxed);
#endif
}

FAV_HD void counter_add(unsigned long long* out, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(out, v);
#else
    reinterpret_cast<std::atomic<unsigned long long>*>(out)->fetch_add(v, std::memory_order_relaxed);
#endif
}

FAV_HD void counter_max(unsigned long long* out, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicMax(out, v);
#else
    auto* a = reinterpret_cast<std::atomic<unsigned long long>*>(out);
    unsigned long long cur = a->load(std::memory_order_relaxed);
    while (cur < v && !a->compare_exchange_weak(cur, v, std::memory_order_relaxed)) {}
#endif
}

template <bool Stats>
struct Counters {
    unsigned long long v[kStatsFields];
    FAV_HD void clear() {
        for (int i = 0; i < kStatsFields; ++i) v[i] = 0;
    }
    FAV_HD void add(int field, unsigned long long n = 1) { v[field] += n; }
    FAV_HD void max(int field, unsigned long long n) { v[field] = n > v[field] ? n : v[field]; }
    // loop passes of this pixel so far: BLA blocks (linear and cubic), exact
    // steps (the plain-double tail's included) and replayed steps
    FAV_HD unsigned long long passes() const {
        return v[ST_LINEAR_BLOCKS] + v[ST_CUBIC_BLOCKS] + v[ST_EXACT] + v[ST_REPLAY];
    }
    // adds the passes to the pixel's work-map entry (none when work is null)
    FAV_HD void flush_work(int32_t* work, size_t index) const {
        if (work) work[index] += static_cast<int32_t>(passes());
    }
    FAV_HD void flush(unsigned long long* out) {
        for (int i = 0; i < kStatsFields; ++i) {
            if (v[i] == 0) continue;
            if (i == ST_MAX_TAIL_STEPS || i == ST_MAX_PIXEL_ITER) counter_max(out + i, v[i]);
            else counter_add(out + i, v[i]);
        }
    }
};

template <>
struct Counters<false> {
    FAV_HD void clear() {}
    FAV_HD void add(int, unsigned long long = 1) {}
    FAV_HD void max(int, unsigned long long) {}
    FAV_HD void flush_work(int32_t*, size_t) const {}
    FAV_HD void flush(unsigned long long*) {}
};

// ---- small exact helpers --------------------------------------------------

// fe_mul(x, 2.0) (renderer.cpp:8268-8270) for the normalised operands of the
// escape margin: the mantissa product x*0.5 lands in [0.25, 0.5), is doubled
// back exactly and the exponent grows by one.
FAV_HD FE fe_times_two(FE x) {
    if (x.m == 0.0) return FE{0.0, 0};
    if (!isfinite_(x.m)) return FE{x.m, 0};
    return FE{x.m, sat_add(x.e, 1)};
}

// |Z_i + d|^2 - R^2 = (|Z_i|^2 - R^2) + 2 Re(Z_i conj d) + |d|^2 with the
// FloatExp parts of Z_i and |Z_i|^2 - R^2 read from the orbit tables
// (sc_escape_margin_with_delta, renderer.cpp:8383-8431).
FAV_HD FE margin_at(const OrbitDev& o, int i, SC delta) {
    const FE zr{o.zr_m[i], o.zr_e[i]}, zi{o.zi_m[i], o.zi_e[i]};
    const FEC d = sc_parts(delta);
    const FE cross = fe_times_two(fe_add(fe_mul(zr, d.re), fe_mul(zi, d.im)));
    const FE dn = fe_add(fe_mul(d.re, d.re), fe_mul(d.im, d.im));
    return fe_add(fe_add(FE{o.margin_m[i], o.margin_e[i]}, cross), dn);
}

FAV_HD SC load_z(const OrbitDev& o, int i) { return SC{o.re[i], o.im[i], o.e[i]}; }

// Z_i + delta, the escape margin and the loop-top norm R^2 + margin
// (renderer.cpp:12915-12923).  When delta is zero or delta.e <= small_exp[i]
// all three equal the orbit-only values (the addends involving delta are
// more than 60 binades below the terms they are added to, where FloatExp
// sums return the larger operand), so they are read from the tables.
FAV_HD void loop_top_values(const OrbitDev& o, int i, SC delta, int mode, SC& total, FE& margin, SN& norm) {
    if (sc_is_zero(delta) || delta.e <= o.small_exp[i]) {
        total = load_z(o, i);
        margin = FE{o.margin_m[i], o.margin_e[i]};
        norm = SN{o.zn_m[i], o.zn_e[i]};
    } else {
        total = sc_add(load_z(o, i), delta);
        margin = margin_at(o, i, delta);
        norm = sn_from_margin(margin, mode);
    }
}

FAV_HD float encode_iteration(int iteration, double bias) {
    return static_cast<float>(static_cast<double>(iteration) - bias);
}

// smooth_escape_value / smooth_escape_scaled, renderer.cpp:8170-8188.  The
// reference takes these logarithms in x87 long double; FP64 differs by at
// most one ulp of the float32 result.
FAV_HD float smooth_escape(int iteration, SN norm, double bias) {
    double value = static_cast<double>(iteration);
    if (norm.m != 0.0) {
        const double lm = 0.5 * (::log(norm.m) + static_cast<double>(norm.e) * kLn2);
        if (lm > 0.0 && isfinite_(lm)) value = static_cast<double>(iteration) - ::log(lm) / kLn2;
    }
    return static_cast<float>(value - bias);
}

struct PlaneOut {
    int64_t* iteration;
    double *phase, *de_x, *de_y, *test1, *test2;
};

FAV_HD void clear_planes(const PlaneOut& p, size_t i, int max_iter) {  // renderer.cpp:8084-8096
    p.iteration[i] = max_iter;
    p.phase[i] = 0.0; p.de_x[i] = 0.0; p.de_y[i] = 0.0;
    p.test1[i] = 0.0; p.test2[i] = 0.0;
}

// render_plane_de, renderer.cpp:7826-7896 (Kalles orientation z*conj(dz)).
FAV_HD bool plane_de(SC total, SC derivative, FE spacing, double& out_re, double& out_im) {
    const double cm = ::hypot(total.re, total.im);
    if (!(cm > 0.0) || !isfinite_(cm)) return false;
    const double lm = ::log(cm) + static_cast<double>(total.e) * kLn2;
    if (!(lm > 0.0) || !isfinite_(lm)) return false;
    const SC unit = sc_from_fe(fe_from_parts(total.re / cm, 0), fe_from_parts(total.im / cm, 0));
    const SC den = sc_mul(unit, sc_mul(sc_conj(derivative), sc_from_fe(spacing, fe_zero())));
    const double den2 = den.re * den.re + den.im * den.im;
    if (!(den2 > 0.0) || !isfinite_(den2)) return false;
    SC r{cm * lm * den.re / den2, -cm * lm * den.im / den2,
         sat_exp(static_cast<long long>(total.e) - den.e)};
    sc_normalize(r);
    out_re = sc_re_double(r);
    out_im = sc_im_double(r);
    return isfinite_(out_re) && isfinite_(out_im);
}

FAV_HD double plane_phase(SC v) {  // renderer.cpp:8075-8082
    const double re = sc_re_double(v), im = sc_im_double(v);
    if (!isfinite_(re) || !isfinite_(im)) return 0.0;
    double phase = ::atan2(im, re) / kTwoPi;
    phase -= ::floor(phase);
    return isfinite_(phase) ? phase : 0.0;
}

FAV_HD void store_escape(const PlaneOut& p, size_t i, int iteration, SC total, SN test1, SN test2,
                         SC derivative, FE spacing) {  // renderer.cpp:8106-8135
    p.iteration[i] = iteration - 1 > 0 ? iteration - 1 : 0;  // Kalles antal = n-1
    p.phase[i] = plane_phase(total);
    double dx = 0.0, dy = 0.0;
    if (!plane_de(total, derivative, spacing, dx, dy)) { dx = 0.0; dy = 0.0; }
    p.de_x[i] = dx;
    p.de_y[i] = dy;
    p.test1[i] = sn_to_double(test1);
    p.test2[i] = sn_to_double(test2);
}

struct Budget {  // renderer.cpp:11872-11890
    const RenderParams* p;
    unsigned ticks;
    FAV_HD bool expired() {
        if (p->budget_ticks == 0) return false;
        if ((ticks++ & 255U) != 0U) return false;
        if (flag_load(p->budget_flag)) return true;
        if (clock_ticks() - *p->budget_start >= p->budget_ticks) {
            flag_store(p->budget_flag);
            return true;
        }
        return false;
    }
};

FAV_HD int highest_level(int max_length) {  // floor(log2(max_length))
    return max_length > 0 ? 31 - __builtin_clz(static_cast<unsigned>(max_length)) : 0;
}

// Wave-level predicates over the active lanes (a single pixel is its own
// "wave" on the host).  Loads indexed by a wave-uniform value take the
// scalar path on the device.
FAV_HD bool wave_all(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(v) == __ballot(1);
#else
    return v;
#endif
}

FAV_HD bool wave_any(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(v) != 0ULL;
#else
    return v;
#endif
}

// Read-only table entry at a wave-uniform index: through the constant
// address space on the device, so the load is a scalar (SMEM) load and the
// branches on its fields are scalar branches.  Tables are written before
// the launch and never during it.
template <typename T>
FAV_HD T load_uniform(const T* table, int64_t index) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const T __attribute__((address_space(4)))* ConstantPtr;
    return ((ConstantPtr)table)[index];
#else
    return table[index];
#endif
}

FAV_HD int wave_first(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// BlaLevels::lookup_linear / lookup, renderer.cpp:8816-8877: the highest
// aligned level whose block is longer than one step and whose validity
// radius exceeds |delta|^2.
template <typename Step>
FAV_HD int lookup_level(const Step* steps, const LevelDir& d, int start, FE dn, int max_length) {
    if (start <= 0 || max_length <= 0 || !fe_finite(dn)) return -1;
    const int base = d.count ? d.size[0] : 0;
    if (start > base) return -1;
    int h = highest_level(max_length);
    if (h > d.count - 1) h = d.count - 1;
    const int offset = start - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= d.size[level]) continue;
        const int64_t k = d.offset[level] + index;
        const Step& s = steps[k];
        const FE r2{s.r2_m, s.r2_e};
        // fe_compare(dn, r2) < 0 with both finite and r2 >= 0
        if (s.length > 1 && fe_finite(r2) && fe_less_nonneg(dn, r2)) return static_cast<int>(k);
    }
    return -1;
}

FAV_HD SC coef(const CubicStepDev& s, int i) { return SC{s.re[i], s.im[i], s.e[i]}; }

// apply_bla_series, renderer.cpp:8686-8731.
FAV_HD SC apply_cubic(const CubicStepDev& s, SC d, SC c, int order) {
    if (order <= 1) return sc_add(sc_mul(coef(s, 0), d), sc_mul(coef(s, 1), c));
    if (order == 2) {
        const SC d_inner = sc_add(sc_mul(coef(s, 2), d), sc_mul(coef(s, 3), c));
        const SC c_inner = sc_add(coef(s, 1), sc_mul(coef(s, 4), c));
        return sc_add(sc_mul(d, sc_add(coef(s, 0), d_inner)), sc_mul(c, c_inner));
    }
    const SC d_cubic = sc_add(sc_add(sc_mul(coef(s, 5), d), sc_mul(coef(s, 6), c)), coef(s, 2));
    const SC c_cubic = sc_add(coef(s, 3), sc_mul(coef(s, 7), c));
    const SC d_inner = sc_add(coef(s, 0), sc_add(sc_mul(d, d_cubic), sc_mul(c, c_cubic)));
    const SC c_inner = sc_add(coef(s, 1), sc_mul(c, sc_add(coef(s, 4), sc_mul(coef(s, 8), c))));
    return sc_add(sc_mul(d, d_inner), sc_mul(c, c_inner));
}

// apply_bla_series_derivative, renderer.cpp:8733-8795.
FAV_HD SC apply_cubic_derivative(const CubicStepDev& s, SC d, SC c, SC der, int order) {
    if (order <= 1) return sc_add(sc_mul(coef(s, 0), der), coef(s, 1));
    SC pd = coef(s, 0), pp = coef(s, 1);
    pd = sc_add(pd, sc_add(sc_double(sc_mul(coef(s, 2), d)), sc_mul(coef(s, 3), c)));
    pp = sc_add(pp, sc_add(sc_mul(coef(s, 3), d), sc_double(sc_mul(coef(s, 4), c))));
    if (order >= 3) {
        const SC d2 = sc_mul(d, d);
        const SC c2 = sc_mul(c, c);
        pd = sc_add(pd, sc_add(sc_mul(sc_add(sc_double(sc_mul(coef(s, 5), d)), sc_mul(coef(s, 5), d)), d),
                               sc_add(sc_double(sc_mul(coef(s, 6), sc_mul(d, c))), sc_mul(coef(s, 7), c2))));
        pp = sc_add(pp, sc_add(sc_mul(coef(s, 6), d2),
                               sc_add(sc_double(sc_mul(coef(s, 7), sc_mul(d, c))),
                                      sc_mul(sc_add(sc_double(sc_mul(coef(s, 8), c)), sc_mul(coef(s, 8), c)), c))));
    }
    return sc_add(sc_mul(pd, der), pp);
}

constexpr SC kOne{1.0, 0.0, 0};

// render_scaled_double_tail, renderer.cpp:12450-12669, split into a
// resumable state and one loop iteration so that a wavefront can run the
// tails of a queue of pixels with lanes refilled as pixels finish.
//
// The launch values the tail reads.  Filled per pixel from that pixel's
// RenderParams (fields of a batch differ).
struct TailParams {
    const double* z_re;
    const double* z_im;
    int32_t orbit_size;
    int32_t max_iter;
    int32_t escape_radius_mode;
    int32_t strict;
    int32_t cycle_enabled;
    int32_t strict_cycle;
    int32_t timed;  // a time budget is set (Budget::expired can return true)
    double output_bias;
    FE pixel_spacing;
};

FAV_HD TailParams tail_params(const RenderParams& p) {
    return TailParams{p.orbit.real_double, p.orbit.imag_double, p.orbit.size, p.max_iter, p.escape_radius_mode,
                      p.strict, p.cycle_enabled, p.strict_cycle, p.budget_ticks != 0 ? 1 : 0, p.output_bias,
                      FE{p.pixel_spacing_m, p.pixel_spacing_e}};
}

// double_tail's loop variables.
struct TailState {
    double dc_re, dc_im, d_re, d_im;
    // Z at ref and at ref + 1, carried from step to step: each step loads
    // Z at ref + 2, so that the load has the whole step to arrive
    double z0_re, z0_im, z1_re, z1_im;
    double tortoise_re, tortoise_im;
    SN previous_norm;
    int cycle_power, cycle_length, tail_steps, tail_iterations, tail_rebases;
    bool cycle_ready;
};

FAV_HD TailState tail_begin(SC dc, SC delta, SN initial_norm) {
    TailState t;
    t.dc_re = sc_re_double(dc);
    t.dc_im = sc_im_double(dc);
    t.d_re = sc_re_double(delta);
    t.d_im = sc_im_double(delta);
    t.tortoise_re = 0.0;
    t.tortoise_im = 0.0;
    t.previous_norm = initial_norm;
    t.cycle_power = 1;
    t.cycle_length = 0;
    t.tail_steps = 0;
    t.tail_iterations = 0;
    t.tail_rebases = 0;
    t.cycle_ready = false;
    return t;
}

// The loop condition of the tail.
FAV_HD bool tail_continues(const TailParams& q, int iteration, int ref) {
    return iteration < q.max_iter && ref >= 0 && ref + 1 < q.orbit_size;
}

// Z at ref and ref + 1 into the carried values (ref + 1 clamped to the
// orbit: a step only runs while ref + 1 < orbit_size).
FAV_HD void tail_load_z(const TailParams& q, int ref, TailState& t) {
    const int r = ref >= 0 && ref < q.orbit_size ? ref : 0;
    const int r1 = r + 1 < q.orbit_size ? r + 1 : r;
    t.z0_re = q.z_re[r]; t.z0_im = q.z_im[r];
    t.z1_re = q.z_re[r1]; t.z1_im = q.z_im[r1];
}

// kTailEscaped: the step escaped; tail_escape writes the result after the
// loop, from the state the step left (iteration, ref, delta, previous norm,
// derivative), so that the lanes of a wave that escape at different steps
// take the smoothing and the plane stores together, once.
enum TailStatus { kTailRunning = 0, kTailEnded = 1, kTailGaveUp = 2, kTailEscaped = 3 };

// One iteration of the tail loop (the caller checked tail_continues).
// kTailEnded / kTailGaveUp end the pixel with `out` (and planes) written as
// double_tail's `return false` / `return true` do.
template <bool Planes, bool Stats>
FAV_HD TailStatus tail_step(const TailParams& q, const PlaneOut& planes, size_t index, int& iteration, int& ref,
                            SC& derivative, Budget& budget, bool& deadline, bool& unresolved, Counters<Stats>& st,
                            TailState& t, float& out) {
    const double escape_squared = escape_r2_double(q.escape_radius_mode);
    // Budget::expired returns false without counting a tick when no budget
    // is set, so skipping the call then changes nothing
    if (q.timed && budget.expired()) {
        out = __builtin_nan(""); deadline = true; unresolved = true;
        st.add(ST_DEADLINE); st.add(ST_UNRESOLVED);
        return kTailEnded;
    }
    const double z_re = t.z0_re, z_im = t.z0_im;
    // Z at ref + 2 for the next step (clamped: ref + 1 < orbit_size here)
    const int ahead = ref + 2 < q.orbit_size ? ref + 2 : ref + 1;
    const double z2_re = q.z_re[ahead], z2_im = q.z_im[ahead];
    if constexpr (Planes) {
        SC prior_total{z_re + t.d_re, z_im + t.d_im, 0};
        sc_normalize(prior_total);
        derivative = sc_add(sc_double(sc_mul(prior_total, derivative)), kOne);
    }
    const double lin_re = 2.0 * (z_re * t.d_re - z_im * t.d_im);
    const double lin_im = 2.0 * (z_re * t.d_im + z_im * t.d_re);
    const double sq_re = t.d_re * t.d_re - t.d_im * t.d_im;
    const double sq_im = 2.0 * t.d_re * t.d_im;
    t.d_re = lin_re + sq_re + t.dc_re;
    t.d_im = lin_im + sq_im + t.dc_im;
    ++ref;
    ++iteration;
    if (++t.tail_iterations > 65536) {
        st.add(ST_TAIL_FALLBACKS); st.add(ST_UNRESOLVED);
        unresolved = true;
        out = q.strict ? __builtin_nan("") : encode_iteration(iteration, q.output_bias);
        return kTailGaveUp;
    }
    st.add(ST_LOGICAL); st.add(ST_EXACT);
    const double t_re = t.z1_re + t.d_re;
    const double t_im = t.z1_im + t.d_im;
    const double m2 = t_re * t_re + t_im * t_im;
    if (!isfinite_(m2)) {
        unresolved = true;
        st.add(ST_GLITCH); st.add(ST_UNRESOLVED);
        out = q.strict ? __builtin_nan("") : encode_iteration(iteration, q.output_bias);
        return kTailEnded;
    }
    if (m2 > escape_squared) return kTailEscaped;
    t.previous_norm = sn_from_double(m2);
    const double dm2 = t.d_re * t.d_re + t.d_im * t.d_im;
    ++t.tail_steps;
    if (q.cycle_enabled && t.tail_steps >= 64 && (t.tail_steps & 31) == 0) {
        if (!t.cycle_ready) {
            t.tortoise_re = t_re; t.tortoise_im = t_im;
            t.cycle_power = 1; t.cycle_length = 0; t.cycle_ready = true;
        } else {
            const double cdr = t_re - t.tortoise_re, cdi = t_im - t.tortoise_im;
            const double cd2 = cdr * cdr + cdi * cdi;
            const int cmin = q.strict_cycle ? 2048 : 512;
            const double ctol = q.strict_cycle ? 1.0e-24 : 1.0e-18;
            if (iteration >= cmin && m2 < 3.0 && cd2 <= ctol * (1.0 < m2 ? m2 : 1.0)) {
                if constexpr (Planes) clear_planes(planes, index, q.max_iter);
                return kTailEnded;
            }
            ++t.cycle_length;
            if (t.cycle_length >= t.cycle_power) {
                t.tortoise_re = t_re; t.tortoise_im = t_im;
                t.cycle_power = t.cycle_power * 2 < (1 << 20) ? t.cycle_power * 2 : (1 << 20);
                t.cycle_length = 0;
            }
        }
    }
    if (m2 < dm2) {
        t.d_re = t_re; t.d_im = t_im;
        ref = 0;
        t.tail_steps = 0;
        t.cycle_ready = false;
        st.add(ST_TAIL_REBASES);
        if (++t.tail_rebases > 64) {
            st.add(ST_TAIL_FALLBACKS); st.add(ST_UNRESOLVED);
            unresolved = true;
            out = q.strict ? __builtin_nan("") : encode_iteration(iteration, q.output_bias);
            if constexpr (Planes) clear_planes(planes, index, q.max_iter);
            return kTailGaveUp;
        }
        tail_load_z(q, 0, t);
        return kTailRunning;
    }
    t.z0_re = t.z1_re; t.z0_im = t.z1_im;
    t.z1_re = z2_re; t.z1_im = z2_im;
    return kTailRunning;
}

// The escape of a step that returned kTailEscaped: |Z+delta|^2 formed again
// from the same operands (the same value), then the smooth value and the
// planes (renderer.cpp:12570-12588).
template <bool Planes>
FAV_HD void tail_escape(const TailParams& q, const PlaneOut& planes, size_t index, int iteration, int ref,
                        const SC& derivative, const TailState& t, float& out) {
    const double escape_squared = escape_r2_double(q.escape_radius_mode);
    const double safe_escape_squared = escape_squared + 1.0e-7;
    const double t_re = q.z_re[ref] + t.d_re;
    const double t_im = q.z_im[ref] + t.d_im;
    const double m2 = t_re * t_re + t_im * t_im;
    const double magnitude = ::sqrt(m2 > safe_escape_squared ? m2 : safe_escape_squared);
    const double qlog = ::log(::log(magnitude)) / kLn2;
    out = static_cast<float>((static_cast<double>(iteration) - qlog) - q.output_bias);
    if constexpr (Planes) {
        SC escaped_total{t_re, t_im, 0};
        sc_normalize(escaped_total);
        store_escape(planes, index, iteration, escaped_total, sn_from_double(m2), t.previous_norm, derivative,
                     q.pixel_spacing);
    }
}

// After the loop condition failed: the reference ran out before the cap.
template <bool Planes, bool Stats>
FAV_HD void tail_exhausted(const TailParams& q, const PlaneOut& planes, size_t index, int iteration, bool& unresolved,
                           Counters<Stats>& st, float& out) {
    if (iteration < q.max_iter) {
        unresolved = true;
        st.add(ST_UNRESOLVED);
        out = q.strict ? __builtin_nan("") : encode_iteration(iteration, q.output_bias);
        if constexpr (Planes) clear_planes(planes, index, q.max_iter);
    }
}

// The whole tail of one pixel.  Returns true when the tail gave up (step
// ceiling or rebase storm).
template <bool Planes, bool Stats>
FAV_HD bool double_tail(const RenderParams& p, const PlaneOut& planes, size_t index, SC dc, int& iteration,
                        int& ref, SC delta, SC& derivative, Budget& budget, bool& deadline, bool& unresolved,
                        Counters<Stats>& st, SN initial_norm, float& out) {
    const TailParams q = tail_params(p);
    TailState t = tail_begin(dc, delta, initial_norm);
    tail_load_z(q, ref, t);
    out = encode_iteration(q.max_iter, q.output_bias);
    TailStatus r = kTailRunning;
    while (tail_continues(q, iteration, ref)) {
        r = tail_step<Planes, Stats>(q, planes, index, iteration, ref, derivative, budget, deadline, unresolved, st,
                                     t, out);
        if (r != kTailRunning) break;
    }
    if (r == kTailEscaped) {
        tail_escape<Planes>(q, planes, index, iteration, ref, derivative, t, out);
        return false;
    }
    if (r != kTailRunning) return r == kTailGaveUp;
    tail_exhausted<Planes, Stats>(q, planes, index, iteration, unresolved, st, out);
    return false;
}

// Direct scaled recurrence after the reference escaped or after a rebase
// storm, renderer.cpp:13044-13128 / 13155-13250.
template <bool Planes, bool Stats>
FAV_HD void exact_tail(const RenderParams& p, const PlaneOut& planes, size_t index, SC dc, int& iteration,
                       SC total, SN total_norm, SC& derivative, Budget& budget, bool& deadline, bool& unresolved,
                       Counters<Stats>& st, float& out) {
    const SC parameter = sc_add(load_z(p.orbit, 1), dc);
    const int margin = p.strict_cycle ? 78 : 60;
    const int cmin = p.strict_cycle ? 2048 : 512;
    SC tortoise = sc_zero();
    int power = 1, length = 0;
    bool ready = false, escaped = false;
    while (iteration < p.max_iter) {
        if (budget.expired()) {
            out = __builtin_nan(""); deadline = true; unresolved = true;
            st.add(ST_DEADLINE); st.add(ST_UNRESOLVED);
            return;
        }
        const SN prior_norm = total_norm;
        if constexpr (Planes) derivative = sc_add(sc_double(sc_mul(total, derivative)), kOne);
        total = sc_add(sc_mul(total, total), parameter);
        ++iteration;
        st.add(ST_LOGICAL); st.add(ST_EXACT);
        total_norm = sc_norm_squared(total);
        if (sc_outside_escape(total_norm, p.escape_radius_mode)) {
            out = smooth_escape(iteration, total_norm, p.output_bias);
            if constexpr (Planes)
                store_escape(planes, index, iteration, total, total_norm, prior_norm, derivative,
                             FE{p.pixel_spacing_m, p.pixel_spacing_e});
            escaped = true;
            break;
        }
        if (p.cycle_enabled && iteration >= cmin && (iteration & 31) == 0
            && sc_compare_norm(total_norm, SN{0.75, 2}) < 0) {
            if (!ready) {
                tortoise = total; power = 1; length = 0; ready = true;
            } else {
                const SN distance = sc_norm_squared(sc_sub(total, tortoise));
                const int scale_exponent = total_norm.e > 0 ? total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale_exponent - margin) {
                    out = encode_iteration(p.max_iter, p.output_bias);
                    escaped = true;
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
    if (!escaped) out = encode_iteration(p.max_iter, p.output_bias);
}

// ---- resumable pixel state ----------------------------------------------
// The loop-top state of one pixel of render_bla_impl (renderer.cpp:12817-
// 12900).  The field is rendered by three kernels that pass this state
// through HBM: a linear-BLA fast kernel, a general kernel with every other
// branch of the loop, and a plain-double tail kernel.  A kernel hands a pixel
// on only at a loop top (or at the tail entry), so the sequence of
// floating-point operations is the one of the single loop.
struct PixelState {
    SC delta, derivative, total;
    SN total_norm, previous_total_norm;
    FE margin;  // margin of (Z[ref], delta) while have_total
    int ref, iteration, rebases;
    bool have_total, pixel_disable_bla;
    unsigned ticks;  // time-budget cadence counter (renderer.cpp:11878-11890)
};

FAV_HD void save_state(const StateStore& q, size_t i, const PixelState& s) {
    q.d_re[i] = s.delta.re; q.d_im[i] = s.delta.im; q.d_e[i] = s.delta.e;
    q.g_re[i] = s.derivative.re; q.g_im[i] = s.derivative.im; q.g_e[i] = s.derivative.e;
    q.t_re[i] = s.total.re; q.t_im[i] = s.total.im; q.t_e[i] = s.total.e;
    q.tn_m[i] = s.total_norm.m; q.tn_e[i] = s.total_norm.e;
    q.pn_m[i] = s.previous_total_norm.m; q.pn_e[i] = s.previous_total_norm.e;
    q.mg_m[i] = s.margin.m; q.mg_e[i] = s.margin.e;
    q.ref[i] = s.ref; q.iter[i] = s.iteration;
    q.misc[i] = (s.rebases & 0xff) | (s.have_total ? 0x100 : 0) | (s.pixel_disable_bla ? 0x200 : 0)
        | static_cast<int>((s.ticks & 0xffU) << 16);
}

FAV_HD PixelState load_state(const StateStore& q, size_t i) {
    PixelState s;
    s.delta = SC{q.d_re[i], q.d_im[i], q.d_e[i]};
    s.derivative = SC{q.g_re[i], q.g_im[i], q.g_e[i]};
    s.total = SC{q.t_re[i], q.t_im[i], q.t_e[i]};
    s.total_norm = SN{q.tn_m[i], q.tn_e[i]};
    s.previous_total_norm = SN{q.pn_m[i], q.pn_e[i]};
    s.margin = FE{q.mg_m[i], q.mg_e[i]};
    s.ref = q.ref[i];
    s.iteration = q.iter[i];
    const int misc = q.misc[i];
    s.rebases = misc & 0xff;
    s.have_total = (misc & 0x100) != 0;
    s.pixel_disable_bla = (misc & 0x200) != 0;
    s.ticks = static_cast<unsigned>(misc >> 16) & 0xffU;
    return s;
}

FAV_HD SC pixel_dc(const RenderParams& p, int px, int py) {  // renderer.cpp:12838-12842
    if (p.point_re) {  // point render: the caller's normalised offsets
        const size_t i = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
        return SC{p.point_re[i], p.point_im[i], p.point_e[i]};
    }
    return sc_from_fe(FE{p.x_off_m[px], p.x_off_e[px]}, FE{p.y_off_m[py], p.y_off_e[py]});
}

FAV_HD PlaneOut plane_out(const RenderParams& p) {
    return PlaneOut{p.plane_iteration, p.plane_phase, p.plane_de_x, p.plane_de_y, p.plane_test1, p.plane_test2};
}

// Pixel set-up and image-series jump, renderer.cpp:12836-12873.
template <bool Planes, bool Stats>
FAV_HD PixelState init_pixel(const RenderParams& p, size_t index, SC dc, Counters<Stats>& st) {
    st.add(ST_PIXELS);
    if constexpr (Planes) clear_planes(plane_out(p), index, p.max_iter);
    PixelState s;
    s.delta = dc;
    s.derivative = kOne;
    s.total = sc_zero();
    s.total_norm = SN{0.0, 0};
    s.previous_total_norm = SN{0.0, 0};
    s.margin = FE{0.0, 0};
    s.ref = 1;
    s.iteration = 1;
    s.rebases = 0;
    s.have_total = false;
    s.pixel_disable_bla = false;
    s.ticks = 0;
    if (p.series.available
        && sc_compare_norm(sc_norm_squared(dc), SN{p.series.radius2_m, p.series.radius2_e}) <= 0) {
        // evaluate_image_series_with_derivative, renderer.cpp:8611-8639
        const int n = p.series.count;
        SC value{p.series.re[n - 1], p.series.im[n - 1], p.series.e[n - 1]};
        SC pd = sc_zero();
        for (int k = n - 1; k > 1; --k) {
            if constexpr (Planes) pd = sc_add(sc_mul(pd, dc), value);
            value = sc_add(sc_mul(value, dc), SC{p.series.re[k - 1], p.series.im[k - 1], p.series.e[k - 1]});
        }
        if constexpr (Planes) s.derivative = sc_add(sc_mul(pd, dc), value);
        s.delta = sc_mul(value, dc);
        s.ref = p.series.iteration;
        s.iteration = p.series.iteration;
        st.add(ST_SERIES_PIXELS);
        st.add(ST_SERIES_JUMPS, static_cast<unsigned long long>(s.iteration));
    }
    return s;
}

enum Outcome : int { kDone = 0, kHandoff = 1, kTail = 2 };

// One wave's derivative log: an entry per block that a loop chose by the
// general rules, (r << 5) | level with r the block's start index (r < 2^24:
// the iteration cap is at most 10^7, renderer.cpp:88-93; level < 32).  The
// plain block loop's common blocks (slot 0 of their rung, run_plain_blocks)
// take no entry: a block that starts at r between two entries is slot 0 of
// rung r, level rungs[r].top, so the replay walks those rungs from one entry
// to the next.  The capacity counts entries; a test build sets a tiny one so
// that replays on a full log run often.
#ifndef FAV_DERIVATIVE_LOG
#define FAV_DERIVATIVE_LOG 1024
#endif
constexpr int kDerivativeLog = FAV_DERIVATIVE_LOG;

FAV_HD int32_t log_entry(int r, int level) { return (r << 5) | level; }

// One block of the derivative update: g <- A g + B (renderer.cpp:13499-13511)
// with the linear step of `level` at index r, linear[dir.offset[level] +
// ((r - 1) >> level)].  The *_nn forms give the general forms' results while
// the operand and the result are normal; otherwise the general forms run.
// Returns the step's length.
FAV_HD int replay_block(const RenderParams& p, int level, int r, SC& g) {
    const LinearStepDev step = load_uniform(p.linear, p.linear_dir.offset[level] + ((r - 1) >> level));
    const SC a{step.a_re, step.a_im, step.a_e};
    const SC b{step.b_re, step.b_im, step.b_e};
    SC g1;
    bool general = !wave_all(sc_normal(g));
    if (!general) {
        g1 = sc_add_nn(sc_mul_nn(a, g), b);
        general = !wave_all(!sc_is_zero(g1) && g1.e <= (1 << 28) && g1.e >= -(1 << 28));
    }
    if (general) g1 = sc_add(sc_mul(a, g), b);
    g = g1;
    return step.length;
}

// The common blocks from r up to index `to`: slot 0 of each rung on the way.
FAV_HD int replay_common(const RenderParams& p, int r, int to, SC& g) {
    while (r < to) {
        const int level = load_uniform(&p.rungs[r].top, 0);
        r += replay_block(p, level, r, g);
    }
    return r;
}

// The derivative after the blocks from reference index `start` to `end`:
// the `count` logged blocks in order, each preceded by the common blocks
// that lead to its start, then the common blocks up to `end`.  Every value
// is equal across the lanes that share the log.
FAV_HD SC replay_derivative(const RenderParams& p, const int32_t* dlog, int count, int start, int end, SC g) {
    count = wave_first(count);
    int r = wave_first(start);
    end = wave_first(end);
    for (int i = 0; i < count; ++i) {
        const int entry = wave_first(dlog[i]);
        r = replay_common(p, r, entry >> 5, g);
        r += replay_block(p, entry & 31, r, g);
    }
    replay_common(p, r, end, g);
    return g;
}

// Logged blocks whose derivative update is still owed: the pixel's
// derivative is its state's derivative advanced through the blocks from
// reference index `start` to `end`, `count` of them logged.
struct DerivativeDebt {
    int count = 0;
    int start = 0;
    int end = 0;
};

template <bool Planes>
FAV_HD void settle_derivative(const RenderParams& p, const int32_t* dlog, PixelState& s, DerivativeDebt& debt) {
    if constexpr (Planes) {
        if (debt.end != debt.start) {
            s.derivative = replay_derivative(p, dlog, debt.count, debt.start, debt.end, s.derivative);
            debt = DerivativeDebt{};
        }
    }
}

// The common iterations of run_linear below, with the same results: while
// the loop top is in the negligible-perturbation regime (Z + delta, the
// margin and the norm are the orbit's own values, so the glitch / escape /
// endpoint tests are the orbit's precomputed flags) and every operand of the
// linear block is a normal ScaledComplex (fav_numeric.h), the block is formed
// with the *_nn operations.  The decisions are taken for the whole wave:
// when every active lane qualifies and all sit at the same reference index
// and choose the same BLA level, the table reads are wave-uniform (scalar)
// loads; any other case returns to run_linear's general loop top, where
// each lane continues from the state of its last committed block.
//
// With planes, the derivative update of each committed block (A g + B,
// renderer.cpp:13499-13511) is deferred: the block's BLA level goes into the
// wave's derivative log, and replay_derivative applies the logged blocks when
// the log is full or when the loop ends with a lane short of the cap.  The
// derivative feeds no decision of the loop, so the replayed value equals the
// one computed block by block, and a wave whose lanes all reach the cap
// (interior pixels) never computes it.  When the loop ends with every lane
// at most one iteration short of the cap, no lane can come back to this loop
// (one more step or block reaches the cap), so the log stays intact and the
// replay is owed in `debt` instead: run_linear settles it only where the
// derivative is read, and not at all for a last step that reaches the cap.
template <bool Planes, bool Stats>
FAV_HD void run_normal_blocks(const RenderParams& p, PixelState& s, SC dc, Counters<Stats>& st, int32_t* dlog,
                              DerivativeDebt& debt) {
    const OrbitDev& o = p.orbit;
    const int max_iter = p.max_iter;
    const LevelDir& dir = p.linear_dir;
    bool ok = s.have_total && s.iteration < max_iter && sc_normal(s.delta);
    int r = wave_first(s.ref);
    if (!wave_all(ok && s.ref == r)) return;
    if (r < 0 || r >= o.size) return;
    // The loop-top record of the wave's reference index.  While blocks
    // commit, the orbit-only loop-top values (total, its norm, the margin)
    // are those of this record; they are written to the pixel state on exit.
    OrbitFastDev top = load_uniform(o.fast, r);
    SC delta = s.delta, derivative = s.derivative;  // the derivative at log_start
    int iteration = s.iteration;
    int log_start = r, logged = 0;
    bool committed = false;
    for (;;) {
        if (o.map_end > 1 && r >= o.map_end) break;
        if ((top.flags & (kOrbitGlitchFlag | kOrbitEscapeFlag)) != 0) break;
        const SN dn = sc_norm_nn(delta);
        const int order = p.approximation_order >= 3 && dn.e < -115 ? (dn.e < -160 ? 1 : 2) : p.approximation_order;
        const int remaining = max_iter - iteration;
        const int limit = p.max_linear_length < remaining ? p.max_linear_length : remaining;
        const int lim = wave_first(limit);
        // the norm compare uses |Z|^2 of the record: equal to the state's
        // total_norm whenever delta is negligible, which the first clause
        // checks; a lane at the cap ends the loop before any further work
        ok = iteration < max_iter && delta.e <= top.small_exp && sn_at_least(SN{top.zn_m, top.zn_e}, dn)
             && order <= 1 && limit == lim;
        if (!wave_all(ok)) break;
        // BLA level: every lane must take the same (highest valid) level
        if (r <= 0 || lim <= 0 || dir.count == 0 || r > dir.size[0]) break;
        int h = highest_level(lim);
        if (h > dir.count - 1) h = dir.count - 1;
        const int offset = r - 1;
        // levels above the alignment of offset are skipped by the lookup
        if (offset != 0 && __builtin_ctz(static_cast<unsigned>(offset)) < h)
            h = __builtin_ctz(static_cast<unsigned>(offset));
        int k = -1, chosen = 0;
        bool mixed = false;
        for (int level = h; level >= 0; --level) {
            const int index = offset >> level;
            if (index >= dir.size[level]) continue;
            const int64_t kk = dir.offset[level] + index;
            const LinearStepDev cand = load_uniform(p.linear, kk);
            if ((cand.pad[0] & kStepUsable) == 0) continue;
            const bool fits = dn.e != cand.r2_e ? dn.e < cand.r2_e : dn.m < cand.r2_m;
            if (wave_all(fits)) {
                k = static_cast<int>(kk);
                chosen = level;
                break;
            }
            if (wave_any(fits)) {  // lanes disagree on the level
                mixed = true;
                break;
            }
        }
        if (k < 0 || mixed) break;
        const LinearStepDev step = load_uniform(p.linear, k);
        const int length = step.length;
        if (r + length > max_iter || (step.pad[0] & kStepNormalCoefficients) == 0) break;
        const int end = r + length;
        const OrbitFastDev next = load_uniform(o.fast, end);
        if ((next.flags & kOrbitEndpointFlag) != 0) break;
        const SC a{step.a_re, step.a_im, step.a_e};
        const SC b{step.b_re, step.b_im, step.b_e};
        const SC d1 = sc_add_nn(sc_mul_nn(a, delta), sc_mul_nn(b, dc));
        // the result stays normal unless the sum cancelled to zero or the
        // exponent left the normal range
        const bool normal = !sc_is_zero(d1) && d1.e <= next.small_exp && d1.e >= -(1 << 28);
        if (!wave_all(normal)) break;
        // commit, renderer.cpp:13499-13511
        delta = d1;
        if constexpr (Planes) {
            if (logged == kDerivativeLog) {
                derivative = replay_derivative(p, dlog, logged, log_start, r, derivative);
                logged = 0;
                log_start = r;
            }
            dlog[logged++] = log_entry(r, chosen);
        }
        iteration += length;
        r = end;
        top = next;
        committed = true;
        st.add(ST_BLA_BLOCKS);
        st.add(ST_LINEAR_BLOCKS);
        st.add(ST_SERIES_JUMPS);
        if constexpr (Stats) {
            int bucket = 0;
            unsigned v = static_cast<unsigned>(length);
            while (v > 1U && bucket < 15) { v >>= 1U; ++bucket; }
            st.add(ST_HISTOGRAM + bucket);
        }
        st.add(ST_LOGICAL, static_cast<unsigned long long>(length));
    }
    if (committed) {
        const SN norm{top.zn_m, top.zn_e};
        if constexpr (Planes) {
            if (logged > 0) {
                if (wave_all(iteration >= max_iter - 1)) {
                    debt.count = logged;
                    debt.start = log_start;
                    debt.end = r;
                } else {
                    derivative = replay_derivative(p, dlog, logged, log_start, r, derivative);
                }
            }
        }
        s.delta = delta;
        s.derivative = derivative;
        s.iteration = iteration;
        s.ref = r;
        s.total = SC{top.z_re, top.z_im, top.z_e};
        s.total_norm = norm;
        s.margin = FE{top.margin_m, top.margin_e};
        s.previous_total_norm = norm;
    }
}

// ---- plain-double block loop ---------------------------------------------
// run_normal_blocks in plain doubles.  Inside the loop every lane holds
// delta = m * 2^E and dc = c * 2^E with one wave-uniform E, and the tables
// hold A, B, the validity radii and |Z|^2 as plain doubles (table_image.cpp
// make_plain_tables).  A block is then m' = A*m + B*c: the reference's
// sc_mul / sc_mul / sc_add (renderer.cpp:7676-7722) without their
// renormalisations, since scaling by a power of two commutes with rounding
// as long as no value is subnormal and the sum is not cut off.  The bounds
// that guarantee this, checked per lane:
//  * m: larger component in [2^-100, 2^100], smaller one 0 or >= 2^-320;
//  * c, A, B: nonzero components in [2^-450, 2^500], the smaller nonzero
//    component at least 2^-420 times the larger;
//  * the two products differ by at most 2^60 in their larger components, so
//    the reference's exponent gap (renderer.cpp:7688) is at most 60 and
//    both are added.
// Then every product, sum and cancellation is a normal double both here and
// in the reference's normalised operands, and m' * 2^E equals the
// reference's delta.  |delta|^2 = |m|^2 * 2^2E is kept a normal double, so
// the loop-top tests and the BLA lookup compare plain doubles with the same
// outcome as the reference's ScaledNorm / FloatExp comparisons.  Any failed
// bound leaves the loop before the block commits; run_normal_blocks then
// continues from the same state.
#ifndef FAV_PLAIN_BLOCK_PROBE
#define FAV_PLAIN_BLOCK_PROBE() ((void)0)
#endif
#ifndef FAV_PLAIN_EXIT_PROBE
#define FAV_PLAIN_EXIT_PROBE(reason) ((void)0)
#endif
constexpr double kPlainLow = 0x1p-100;
constexpr double kPlainHigh = 0x1p100;
constexpr double kPlainTiny = 0x1p-320;
constexpr double kOperandLow = 0x1p-450;
constexpr double kOperandHigh = 0x1p500;
constexpr double kOperandSpan = 0x1p-420;

FAV_HD double min_abs(double x, double y) { return __builtin_fmin(fabs_(x), fabs_(y)); }

// the scaled delta's bounds
FAV_HD bool plain_delta_ok(double re, double im) {
    const double hi = max_abs(re, im), lo = min_abs(re, im);
    return hi >= kPlainLow && hi <= kPlainHigh && (lo == 0.0 || lo >= kPlainTiny);
}

// the operand bounds of a value whose span is already known to be within
// them (a rescaled delta or dc)
FAV_HD bool plain_value_ok(double re, double im) {
    const double hi = max_abs(re, im), lo = min_abs(re, im);
    return hi >= kOperandLow && hi <= kOperandHigh && (lo == 0.0 || lo >= kOperandLow);
}

// dc's bounds, from its normalised mantissas (the span test runs on them,
// so a component that scaling would flush to zero is caught) and its
// scaled value
FAV_HD bool plain_operand_ok(SC v, double re, double im) {
    const double span_hi = max_abs(v.re, v.im), span_lo = min_abs(v.re, v.im);
    const double hi = max_abs(re, im), lo = min_abs(re, im);
    return (span_lo == 0.0 || span_lo >= span_hi * kOperandSpan) && hi >= kOperandLow && hi <= kOperandHigh
        && (lo == 0.0 || lo >= kOperandLow);
}

// Ballot of `v` over the active lanes, and the active lanes themselves (a
// single pixel is its own one-lane wave on the host).
FAV_HD uint64_t wave_mask(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(v);
#else
    return v ? 1ULL : 0ULL;
#endif
}

FAV_HD uint64_t wave_active() {
#if defined(__HIP_DEVICE_COMPILE__)
    // a ballot (uniform to the compiler's divergence analysis) rather than a
    // read of the exec mask, so that tests against it are uniform branches
    return __builtin_amdgcn_ballot_w64(true);
#else
    return 1ULL;
#endif
}

enum PlainExit : int {
    kPlainNotEntered = 0,  // the state is unchanged
    kPlainLoopTop = 1,     // the state is a loop top at s.ref with a negligible delta
    kPlainInterior = 2,    // ... whose step reaches the cap: the pixel is interior
};

// A wave-uniform double kept in scalar registers: readfirstlane of both
// halves tells the compiler the value is uniform (a no-op on a value that
// is already scalar).
FAV_HD double uniform_double(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b & 0xffffffffULL)));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32)));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
#else
    return v;
#endif
}

// A step read field by field through scalar loads.
FAV_HD PlainStepDev load_step(const PlainStepDev* t) {
    return PlainStepDev{uniform_double(load_uniform(&t->a_re, 0)), uniform_double(load_uniform(&t->a_im, 0)),
                        uniform_double(load_uniform(&t->b_re, 0)), uniform_double(load_uniform(&t->b_im, 0))};
}

// A uniform radius slot: fits (no lane at or above it) and applicable (sign
// bit clear), read as an integer test of the high word.  On the device the
// word goes through readfirstlane, so that the test stays a 32-bit scalar
// compare instead of being folded into a 64-bit vector one.
FAV_HD bool slot_plain(double t) {
    const int hi = static_cast<int>(__builtin_bit_cast(unsigned long long, t) >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(hi) >= 0;
#else
    return hi >= 0;
#endif
}

// What a slot-0 block reads: the first cache line of its rung.
struct RungLine {
    int32_t end_hi, gap_lo_end, gap_hi_end;
    PlainStepDev step;
    uint32_t stride;
};

FAV_HD RungLine load_line(const RungDev* g) {
    return RungLine{load_uniform(&g->end_hi, 0), load_uniform(&g->gap_lo_end, 0), load_uniform(&g->gap_hi_end, 0),
                    PlainStepDev{uniform_double(load_uniform(&g->a_re, 0)), uniform_double(load_uniform(&g->a_im, 0)),
                                 uniform_double(load_uniform(&g->b_re, 0)), uniform_double(load_uniform(&g->b_im, 0))},
                    load_uniform(&g->stride, 0)};
}

// Marks a scalar value as needed here, so that the load that produces it is
// issued with the loads before it rather than sunk to its first use (an
// empty asm statement: no instruction is emitted).
FAV_HD void issue_here(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : : "s"(v));
#else
    (void)v;
#endif
}

// The integer fields of a rung line read here rather than where they are
// used (see issue_here).
FAV_HD void issue_line(const RungLine& line) {
    issue_here(line.stride);
    issue_here(static_cast<uint32_t>(line.end_hi));
    issue_here(static_cast<uint32_t>(line.gap_lo_end));
    issue_here(static_cast<uint32_t>(line.gap_hi_end));
}

// Keeps the instructions before it (a rung read issued a block ahead) from
// being scheduled after the ones that follow it (no instruction is
// emitted).
FAV_HD void schedule_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// The step of slot k of a rung.
FAV_HD PlainStepDev slot_step(const RungDev* g, int k) {
    if (k == 0)
        return PlainStepDev{uniform_double(load_uniform(&g->a_re, 0)), uniform_double(load_uniform(&g->a_im, 0)),
                            uniform_double(load_uniform(&g->b_re, 0)), uniform_double(load_uniform(&g->b_im, 0))};
    return load_step(&g->step_low[k - 1]);
}

// The scalings a run of slot-0 blocks applies to the rung bounds (fav_device.h
// "hi" words), fixed while E and c are: as integers subtracted from a high
// word they multiply the bound by a power of two, exactly.
//  * shift_hi = 2E << 20: end_hi bounds |delta'|^2 = |m'|^2 2^2E, so
//    end_hi - shift_hi bounds |m'|^2;
//  * lo_shift = fl_min << 20 and hi_shift = (fl_max + 1) << 20, with
//    2^fl_min <= 1/|c|^2 < 2^(fl_max + 1) for every lane (1/|c|^2 being the
//    lane's inv_cc): |m'|^2 >= gap_lo 2^-fl_min implies |m'|^2 inv_cc >=
//    gap_lo, and |m'|^2 <= gap_hi 2^-(fl_max + 1) implies |m'|^2 inv_cc <=
//    gap_hi, which are the product-gap tests on the next block's |m|^2.
// The ranges (-2E in [0, 820], fl in [-256, 256], hi words of bounds in
// [2^-900, 2^600] or -0.0) keep every difference a finite double's high
// word of the same sign, so no subtraction overflows.
struct RunScale {
    int32_t shift_hi, lo_shift, hi_shift;
};

// A double from its high word (low word 0).
FAV_HD double from_high_word(int32_t hi) {
    return __builtin_bit_cast(double, static_cast<unsigned long long>(static_cast<uint32_t>(hi)) << 32);
}

// The largest (want_max) or smallest value of v over the active lanes,
// every v in [lo, hi]: bisection on ballots, so the result is wave-uniform
// and lanes outside the exec mask take no part.
FAV_HD int wave_extreme(int v, int lo, int hi, bool want_max) {
#if defined(__HIP_DEVICE_COMPILE__)
    int a = lo, b = hi;  // the extreme is in [a, b]
    while (a < b) {
        if (want_max) {
            const int mid = a + (b - a + 1) / 2;
            if (__ballot(v >= mid) != 0ULL) a = mid; else b = mid - 1;
        } else {
            const int mid = a + (b - a) / 2;
            if (__ballot(v <= mid) != 0ULL) b = mid; else a = mid + 1;
        }
    }
    return a;
#else
    (void)lo; (void)hi; (void)want_max;
    return v;
#endif
}

// The gap scalings of RunScale for the lanes' inv_cc (floor(log2 inv_cc),
// 300 for a value that is not a positive finite double); false when some
// lane's is outside [-256, 255], where the bounds would leave their range.
FAV_HD bool gap_scale(double inv_cc, RunScale& k) {
    int fl = 300;
    if (inv_cc > 0.0 && isfinite_(inv_cc)) {
        fl = frexp_exp_(inv_cc) - 1;
        fl = fl < -300 ? -300 : (fl > 300 ? 300 : fl);
    }
    const int fl_min = wave_extreme(fl, -300, 300, false);
    const int fl_max = wave_extreme(fl, -300, 300, true);
    k.lo_shift = fl_min * (1 << 20);
    k.hi_shift = (fl_max + 1) * (1 << 20);
    return fl_min >= -256 && fl_max <= 255;
}

// min(a, b) of two wave-uniform values on the scalar unit (left to itself
// the compiler forms the bound's three-way minimum on the vector unit).
FAV_HD int32_t scalar_min(int32_t a, int32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    int32_t r;
    __asm__("s_min_i32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b) : "scc");
    return r;
#else
    return a < b ? a : b;
#endif
}

// The upper bound on |m'|^2 of a slot-0 block: the end index's bound on
// |delta'|^2 and its rung's gap_hi, both scaled, and 2^199 (so that the
// larger component of m' stays below 2^100), by one integer minimum each;
// -0.0 (negative high word) stays negative, so that no |m'|^2 is below it.
FAV_HD double upper_bound(const RungLine& line, const RunScale& k) {
    constexpr int32_t kCapHigh = 0x4c600000;  // high word of 2^199
    return from_high_word(scalar_min(scalar_min(line.end_hi - k.shift_hi, line.gap_hi_end - k.hi_shift), kCapHigh));
}

// The lower bound on |m'|^2: the end index's rung's gap_lo, scaled.
FAV_HD double lower_bound(const RungLine& line, const RunScale& k) {
    return from_high_word(line.gap_lo_end - k.lo_shift);
}

// One block m' = A*m + B*c and the tests of its result, taken together:
// both products added (exponent gap at most 60), m's bounds, and negligible
// at the end index (the condition of the reference's loop-top values there
// being the orbit's).
struct PlainBlock {
    double nr, ni, nx;
    uint64_t added, bounded, negligible;
};

FAV_HD PlainBlock plain_block(const PlainStepDev& step, double mr, double mi, double cr, double ci, int E,
                              double next_small) {
    // A*delta + B*dc, each product in sc_mul's operation order
    const double pr = step.a_re * mr - step.a_im * mi;
    const double pi = step.a_re * mi + step.a_im * mr;
    const double qr = step.b_re * cr - step.b_im * ci;
    const double qi = step.b_re * ci + step.b_im * cr;
    PlainBlock b;
    b.nr = pr + qr;
    b.ni = pi + qi;
    const double pm = max_abs(pr, pi), qm = max_abs(qr, qi);
    b.added = wave_mask(__builtin_fmax(pm, qm) <= __builtin_fmin(pm, qm) * 0x1p60);
    b.nx = max_abs(b.nr, b.ni);
    const double nn = min_abs(b.nr, b.ni);
    b.bounded = wave_mask(b.nx >= kPlainLow) & wave_mask(b.nx <= kPlainHigh)
        & (wave_mask(nn == 0.0) | wave_mask(nn >= kPlainTiny));
    b.negligible = wave_mask(ldexp_(b.nx, E) < next_small);
    return b;
}

// A slot-0 block and its tests as one ballot, through the rung's bounds
// (fav_device.h), each of which implies the corresponding general test of
// plain_block or of the loop top (so a block this accepts is the block the
// general rules commit, and one it rejects is decided by them).  The tests
// on a block's input, m, are made on the result of the block before it
// (for the first block of a run, by the caller, first_block_fits):
//  * |delta|^2 < fast_r2 at the block's rung: the loop-top tests and slot 0
//    chosen;
//  * gap_lo <= |m|^2 / |c|^2 <= gap_hi: |A m| / |B c| within 2^+-58 up to
//    rounding, so the larger components are within 2^59 < 2^60;
// and on its result, m':
//  * both components of m' at least 2^-99 (so the larger one is at least
//    2^-100 and the smaller one at least 2^-320; a component closer to
//    zero is left to the general test) and |m'|^2 < 2^199 (the larger
//    component at most 2^100);
//  * |delta'|^2 below min(small^2, fast_r2 of the end index's rung), both
//    with a factor (1 - 2^-48): the larger component of delta' below small
//    (|m'|^2 carries a relative error below 2^-50), and the next block's
//    loop-top bound;
//  * the next block's product gap, through the lane-independent bounds of
//    RunScale.
// The three bounds on |m'|^2 are two compares with scalar operands
// (upper_bound, lower_bound).  |m'|^2 is the next block's |m|^2, formed as
// fma(nr, nr, ni * ni): within 2^-50 of the reference's nr * nr + ni * ni,
// which the (1 - 2^-48) factors and the gap margins absorb; the committed
// m' itself is the reference's value.
struct CommonBlock {
    double nr, ni, sq;
    uint64_t ok;
};

FAV_HD CommonBlock common_block(const RungLine& line, double mr, double mi, double cr, double ci,
                                const RunScale& k) {
    const double pr = line.step.a_re * mr - line.step.a_im * mi;
    const double pi = line.step.a_re * mi + line.step.a_im * mr;
    const double qr = line.step.b_re * cr - line.step.b_im * ci;
    const double qi = line.step.b_re * ci + line.step.b_im * cr;
    CommonBlock b;
    b.nr = pr + qr;
    b.ni = pi + qi;
    b.sq = __builtin_fma(b.nr, b.nr, b.ni * b.ni);
    const double nn = min_abs(b.nr, b.ni);
    b.ok = wave_mask(nn >= 0x1p-99) & wave_mask(b.sq < upper_bound(line, k))
        & wave_mask(b.sq >= lower_bound(line, k));
    return b;
}

// The counters of a committed slot-0 block of byte stride `stride`.
template <bool Stats>
FAV_HD void plain_common_stats(Counters<Stats>& st, uint32_t stride) {
    FAV_PLAIN_BLOCK_PROBE();
    st.add(ST_BLA_BLOCKS);
    st.add(ST_LINEAR_BLOCKS);
    st.add(ST_SERIES_JUMPS);
    if constexpr (Stats) {
        const int level = __builtin_ctz(stride) - 8;
        st.add(ST_HISTOGRAM + (level < 15 ? level : 15));
    }
    st.add(ST_LOGICAL, static_cast<unsigned long long>(stride >> 8));
}

template <bool Planes, bool Stats>
FAV_HD PlainExit run_plain_blocks(const RenderParams& p, PixelState& s, SC dc, Counters<Stats>& st, int32_t* dlog,
                                  DerivativeDebt& debt) {
    FAV_PLAIN_EXIT_PROBE(14);
    const int max_iter = p.max_iter;
    const RungDev* rungs = p.rungs;
    if (rungs == nullptr || p.linear_dir.count == 0 || p.approximation_order == 2) return kPlainNotEntered;
    const int r0 = wave_first(s.ref);
    const int it0 = wave_first(s.iteration);
    const bool entry = s.iteration < max_iter && sc_normal(s.delta) && s.ref == r0 && s.iteration == it0;
    if (!wave_all(entry)) { FAV_PLAIN_EXIT_PROBE(0); return kPlainNotEntered; }
    if (r0 <= 0 || r0 >= p.orbit.size) return kPlainNotEntered;
    int E = wave_first(s.delta.e);
    // |delta|^2 = |m|^2 * 2^2E with |m| >= 2^-100 stays a normal double
    // while E >= -410; |delta| < |Z| keeps it finite.  E <= 0 (a negligible
    // delta is far below 1) bounds the scaling of end_bound.
    if (E < -410 || E > 0) return kPlainNotEntered;
    // the rungs start at the table's level cap, which must bound the
    // lookup's highest level (floor(log2) of the length limit)
    const int max_length = p.max_linear_length;
    const int full_level = highest_level(max_length);
    if (full_level > p.plain_top) return kPlainNotEntered;
    double mr = ldexp_(s.delta.re, s.delta.e - E), mi = ldexp_(s.delta.im, s.delta.e - E);
    double cr = ldexp_(dc.re, dc.e - E), ci = ldexp_(dc.im, dc.e - E);
    const double mx = max_abs(mr, mi);
    // delta's span from its mantissas, as for dc: 2^-220 keeps the smaller
    // component of m at or above 2^-320.  Delta must be negligible at r0
    // (delta.e <= small_exp: the larger component below 2^small_exp;
    // small_d = 0 also stands for the glitch / escape flags and the
    // reference's escape), as it is at every index a block ends on, so that
    // the loop-top values at the exit index are the orbit's own.
    const double dspan_hi = max_abs(s.delta.re, s.delta.im), dspan_lo = min_abs(s.delta.re, s.delta.im);
    if (!wave_all(plain_delta_ok(mr, mi) && (dspan_lo == 0.0 || dspan_lo >= dspan_hi * 0x1p-220)
                  && plain_operand_ok(dc, cr, ci) && ldexp_(mx, E) < load_uniform(&rungs[r0].small_d, 0))) {
        FAV_PLAIN_EXIT_PROBE(1);
        return kPlainNotEntered;
    }
    // From here every exit writes the whole loop-top state (delta, the
    // derivative, the indices and the orbit's loop-top values at r), so
    // nothing of the entry state stays live across the loop.
    // Order 1 (renderer.cpp:13252-13273): |delta|^2 < 2^-161 when the
    // approximation order is 3; order 2 never takes a linear block.
    const double order_bound = p.approximation_order >= 3 ? 0x1p-161 : __builtin_inf();
    // 1 / |c|^2, for the slot-0 blocks' product-gap test, and its wave-wide
    // scalings (gap_ok false: no slot-0 run at this scale)
    double inv_cc = 1.0 / (cr * cr + ci * ci);
    RunScale scale{};
    bool gap_ok = gap_scale(inv_cc, scale);
    // iteration - r stays `lag`: both advance by each block's length
    const int lag = it0 - r0;
    // below this iteration the length limit is max_length itself
    const int last_full = max_iter - max_length;
    // r is kept as the byte offset of its rung (r < 2^24, so it fits)
    const char* const base = reinterpret_cast<const char*>(rungs);
    uint32_t off = static_cast<uint32_t>(r0) << 8;
    // A common block starts at r <= r_full (iteration <= last_full: the
    // length limit is max_length, so the lookup starts at the top level,
    // and r <= iteration keeps the block below the cap); r >= 1 here
    const int r_full = last_full - lag;
    const uint32_t off_full = r_full > 0 ? static_cast<uint32_t>(r_full) << 8 : 0u;
    double sq = mr * mr + mi * mi;
    SC derivative = s.derivative;  // the derivative at log_start
    int log_start = r0, logged = 0;
    bool committed = false;
    // The block loop runs at a fixed scale E.  A result outside m's bounds
    // ends it without a commit; the committed delta and dc are then rescaled
    // by the result's magnitude (first lane) and the block is tried again.
    // A rescaled m only needs the operand bounds (its span is unchanged),
    // and |delta|^2 is unchanged, so every bound above still holds.
    bool retried = false;
    for (;;) {
        int shift = 0;
        bool rescale = false;
        for (;;) {
            const uint64_t active = wave_active();
            // Common blocks: slot 0 of the rung while every lane's |delta|^2
            // is below its fast bound (every loop-top test of the block
            // passes, fav_device.h) and r <= r_full.  They take no log entry.
            // small_d > 0 at r (checked at entry and at every commit) means a
            // block may start here; the other waves of the CU walk the same
            // rungs.
            // Slot-0 blocks (common_block), one rung read per block, issued a
            // block ahead: while a block is formed and tested, the rung its
            // end index names is already on its way.  That read is in the
            // table whether or not the block commits: a rung whose slot 0
            // applies has its end index in the table, and one whose slot 0
            // does not has stride 0 (it names itself).
            const uint32_t off_common = off;
            // the first block's tests on its input (the loop-top bound and
            // the product gap; the later blocks' are made by the block
            // before them)
            const RungDev* const g0 = reinterpret_cast<const RungDev*>(base + off);
            scale.shift_hi = 2 * E * (1 << 20);
            if (gap_ok
                && (wave_mask(ldexp_(sq, 2 * E) < uniform_double(load_uniform(&g0->fast_r2, 0)))
                    & wave_mask(sq >= from_high_word(load_uniform(&g0->gap_lo_own, 0) - scale.lo_shift))
                    & wave_mask(sq <= from_high_word(load_uniform(&g0->gap_hi_own, 0) - scale.hi_shift)))
                    == active) {
                RungLine line = load_line(g0);
                // unrolled twice, so that m and the rung lines alternate
                // between two sets of registers instead of being copied
                for (;;) {
                    const uint32_t off2 = off + line.stride;
                    const RungLine line2 = load_line(reinterpret_cast<const RungDev*>(base + off2));
                    schedule_fence();
                    const CommonBlock b = common_block(line, mr, mi, cr, ci, scale);
                    // the next address needs line2's stride: read it here,
                    // after the block, so that it is not left to the commit
                    schedule_fence();
                    issue_line(line2);
                    if ((b.ok & (off <= off_full ? ~0ULL : 0ULL)) != active) break;
                    // commit, renderer.cpp:13499-13511
                    plain_common_stats(st, line.stride);
                    off = off2;
                    line = load_line(reinterpret_cast<const RungDev*>(base + off2 + line2.stride));
                    schedule_fence();
                    const CommonBlock b2 = common_block(line2, b.nr, b.ni, cr, ci, scale);
                    schedule_fence();
                    issue_line(line);
                    if ((b2.ok & (off <= off_full ? ~0ULL : 0ULL)) != active) {
                        mr = b.nr;
                        mi = b.ni;
                        break;
                    }
                    mr = b2.nr;
                    mi = b2.ni;
                    plain_common_stats(st, line2.stride);
                    off += line2.stride;
                }
            }
            if (off != off_common) {
                committed = true;
                retried = false;
                sq = mr * mr + mi * mi;
            }
            // Any other block, by the lookup of run_normal_blocks: the
            // loop-top tests that depend on |delta|^2 (|Z|^2 >= |delta|^2 and
            // order 1), then the highest aligned level at or below log2 of
            // the limit whose radius every lane's |delta|^2 is below; slots
            // above the limit are skipped, and the slots end at level 1.
            const int r = static_cast<int>(off >> 8);
            const int iteration = r + lag;
            if (iteration >= max_iter) { FAV_PLAIN_EXIT_PROBE(2); break; }
            const RungDev* g = rungs + r;
            const double dd = ldexp_(sq, 2 * E);
            const double zn_d = uniform_double(load_uniform(&g->zn_d, 0));
            if ((wave_mask(dd <= zn_d) & wave_mask(dd < order_bound)) != active) { FAV_PLAIN_EXIT_PROBE(4); break; }
            const int top_level = load_uniform(&g->top, 0);
            int limit_level = full_level;
            if (iteration > last_full) limit_level = highest_level(max_iter - iteration);
            const int first = top_level > limit_level ? top_level - limit_level : 0;
            const int slots = top_level < kRungSlots ? top_level : kRungSlots;
            int slot = -1;
            for (int k = first; k < slots; ++k) {
                const double r2 = uniform_double(load_uniform(&g->r2[k], 0));
                const uint64_t fits = wave_mask(dd < fabs_(r2));
                if (fits == 0) continue;
                // else: lanes disagree, or the step is not plain
                if (fits == active && slot_plain(r2)) slot = k;
                break;
            }
            if (slot < 0) { FAV_PLAIN_EXIT_PROBE(9); break; }
            const int level = top_level - slot;
            const int end = r + (1 << level);
            if (end > max_iter) { FAV_PLAIN_EXIT_PROBE(10); break; }
            const double next_small = uniform_double(load_uniform(&rungs[end].small_d, 0));
            const PlainBlock b = plain_block(slot_step(g, slot), mr, mi, cr, ci, E, next_small);
            if ((b.added & b.bounded & b.negligible) != active) {
                if (b.added != active) {
                    FAV_PLAIN_EXIT_PROBE(12);
                } else if (b.bounded != active) {
                    shift = wave_first(frexp_exp_nz(b.nx));
                    rescale = true;
                } else {
                    FAV_PLAIN_EXIT_PROBE(11);
                }
                break;
            }
            // commit, renderer.cpp:13499-13511, with a log entry
            mr = b.nr;
            mi = b.ni;
            sq = mr * mr + mi * mi;
            if constexpr (Planes) {
                if (logged == kDerivativeLog) {
                    derivative = replay_derivative(p, dlog, logged, log_start, r, derivative);
                    logged = 0;
                    log_start = r;
                }
                dlog[logged++] = log_entry(r, level);
            }
            off = static_cast<uint32_t>(end) << 8;
            committed = true;
            retried = false;
            FAV_PLAIN_BLOCK_PROBE();
            st.add(ST_BLA_BLOCKS);
            st.add(ST_LINEAR_BLOCKS);
            st.add(ST_SERIES_JUMPS);
            if constexpr (Stats) st.add(ST_HISTOGRAM + (level < 15 ? level : 15));
            st.add(ST_LOGICAL, static_cast<unsigned long long>(1 << level));
        }
        if (!rescale || retried || E + shift < -410 || E + shift > 0) {
            if (rescale) FAV_PLAIN_EXIT_PROBE(13);
            break;
        }
        const double sr = ldexp_(mr, -shift), si = ldexp_(mi, -shift);
        const double tr = ldexp_(cr, -shift), ti = ldexp_(ci, -shift);
        if (!wave_all(plain_value_ok(sr, si) && plain_value_ok(tr, ti))) {
            FAV_PLAIN_EXIT_PROBE(13);
            break;
        }
        mr = sr;
        mi = si;
        cr = tr;
        ci = ti;
        E += shift;
        sq = mr * mr + mi * mi;
        inv_cc = 1.0 / (cr * cr + ci * ci);
        gap_ok = gap_scale(inv_cc, scale);
        retried = true;
    }
    // At one iteration below the cap no block fits (lengths are at least 2),
    // so the loop top takes an exact step to the cap, and the pixel is
    // interior, when none of its tests fires: no glitch / escape / map end
    // (small_d > 0, as checked for r), no rebase (|Z|^2 >= |delta|^2), order
    // 1 and no plain-double tail (|delta|^2 < 2^-90; renderer.cpp:13252-
    // 13290).  That pixel reads no derivative.
    const int r = static_cast<int>(off >> 8);
    const int iteration = r + lag;
    const double dd_end = ldexp_(sq, 2 * E);
    const bool interior = iteration == max_iter - 1
        && wave_all(dd_end <= load_uniform(&rungs[r].zn_d, 0) && dd_end < order_bound && dd_end < 0x1p-90);
    if constexpr (Planes) {
        if (r != log_start && !interior) {
            if (wave_all(iteration >= max_iter - 1)) {
                debt.count = logged;
                debt.start = log_start;
                debt.end = r;
            } else {
                derivative = replay_derivative(p, dlog, logged, log_start, r, derivative);
            }
        }
    }
    // the loop-top state at r: delta is negligible there, so Z + delta, its
    // norm and the margin are the orbit's own values (the state a committed
    // block leaves, renderer.cpp:13499-13511, with the previous norm equal
    // to the new one); without a commit they are the values the loop top
    // at r0 computes, and the previous norm stays
    const OrbitFastDev top = load_uniform(p.orbit.fast, r);
    const SN norm{top.zn_m, top.zn_e};
    const int k = frexp_exp_nz(max_abs(mr, mi));
    s.delta = SC{ldexp_(mr, -k), ldexp_(mi, -k), E + k};
    s.derivative = derivative;
    s.iteration = iteration;
    s.ref = r;
    s.total = SC{top.z_re, top.z_im, top.z_e};
    s.total_norm = norm;
    s.margin = FE{top.margin_m, top.margin_e};
    if (committed) s.previous_total_norm = norm;
    s.have_total = true;
    return interior ? kPlainInterior : kPlainLoopTop;
}

// ---- the pixel on the reference --------------------------------------------
// lookup_level (renderer.cpp:8816-8877) with |delta|^2 = 0, through scalar
// loads of the fields it reads: the highest aligned level whose step is
// longer than one step and has a finite, nonzero radius.
template <typename Step>
FAV_HD int lookup_level_zero(const Step* steps, const LevelDir& d, int start, int max_length) {
    if (start <= 0 || max_length <= 0 || d.count == 0 || start > d.size[0]) return -1;
    int h = highest_level(max_length);
    if (h > d.count - 1) h = d.count - 1;
    const int offset = start - 1;
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= d.size[level]) continue;
        const int64_t k = d.offset[level] + index;
        const int length = load_uniform(&steps[k].length, 0);
        const double r2 = load_uniform(&steps[k].r2_m, 0);
        if (length > 1 && isfinite_(r2) && r2 != 0.0) return static_cast<int>(k);
    }
    return -1;
}

// The loop of linear_loop for a pixel sitting on the reference (delta = dc
// = 0): Z + delta is Z, so every loop top takes the orbit's own values (the
// glitch and escape tests are the orbit's flags, |delta|^2 = 0 never
// rebases), the lookup runs with |delta|^2 = 0 in the table of the
// approximation order (order 1: linear; otherwise the cubic table, whose
// block is A*0 + B*0 = 0 as in linear_loop), and delta stays zero.  Table
// reads are scalar; any other branch returns to linear_loop at the same loop
// top.
template <bool Planes, bool Stats>
FAV_HD void run_reference_blocks(const RenderParams& p, PixelState& s, Counters<Stats>& st) {
    const OrbitDev& o = p.orbit;
    const int max_iter = p.max_iter;
    int r = wave_first(s.ref);
    int iteration = wave_first(s.iteration);
    if (!wave_all(s.ref == r && s.iteration == iteration && sc_is_zero(s.delta))) return;
    const bool linear = p.approximation_order <= 1;
    SC derivative = s.derivative;
    bool committed = false;
    // The walk's record at r (p.ref_walk), read a block ahead: while the
    // limit min(max length, remaining) has the walk's floor(log2), a record
    // with length > 0 is the block the lookup below commits at r.
    const RefWalkDev* const walk = p.ref_walk;
    RefWalkDev w{};
    if (walk != nullptr && r >= 0 && r < o.size) w = load_uniform(walk, r);
    for (;;) {
        if (iteration >= max_iter || r < 0 || r >= o.size) break;
        const int remaining = max_iter - iteration;
        int length;
        SC a, b;
        if (walk != nullptr && w.length > 0 && remaining >= p.ref_walk_limit) {
            length = w.length;
            a = SC{w.a_re, w.a_im, w.a_e};
            b = SC{w.b_re, w.b_im, w.b_e};
            if (r + length > max_iter) break;
        } else {
            if ((load_uniform(&o.fast[r].flags, 0) & (kOrbitGlitchFlag | kOrbitEscapeFlag)) != 0) break;
            if (o.map_end > 1 && r >= o.map_end) break;
            int k;
            if (linear) {
                k = lookup_level_zero(p.linear, p.linear_dir, r,
                                      p.max_linear_length < remaining ? p.max_linear_length : remaining);
                if (k < 0) break;
                const LinearStepDev step = load_uniform(p.linear, k);
                length = step.length;
                a = SC{step.a_re, step.a_im, step.a_e};
                b = SC{step.b_re, step.b_im, step.b_e};
            } else {
                k = lookup_level_zero(p.cubic, p.cubic_dir, r,
                                      p.max_bla_length < remaining ? p.max_bla_length : remaining);
                if (k < 0) break;
                const CubicStepDev* step = p.cubic + k;
                length = load_uniform(&step->length, 0);
                if constexpr (Planes) {
                    a = SC{load_uniform(&step->re[0], 0), load_uniform(&step->im[0], 0),
                           load_uniform(&step->e[0], 0)};
                    b = SC{load_uniform(&step->re[1], 0), load_uniform(&step->im[1], 0),
                           load_uniform(&step->e[1], 0)};
                }
            }
            if (r + length > max_iter) break;
            // the endpoint's loop-top norm is |Z|^2 there (renderer.cpp:13422)
            if ((load_uniform(&o.fast[r + length].flags, 0) & kOrbitEndpointFlag) != 0) break;
        }
        const int end = r + length;
        if (walk != nullptr) {
            // (an end past the orbit ends the walk at the loop top)
            w = load_uniform(walk, end < o.size ? end : r);
            schedule_fence();
        }
        if constexpr (Planes) derivative = sc_add(sc_mul(a, derivative), b);
        // commit, renderer.cpp:13499-13511
        r = end;
        iteration += length;
        committed = true;
        st.add(ST_BLA_BLOCKS);
        st.add(linear ? ST_LINEAR_BLOCKS : ST_CUBIC_BLOCKS);
        st.add(ST_SERIES_JUMPS);
        if constexpr (Stats) {
            int bucket = 0;
            unsigned v = static_cast<unsigned>(length);
            while (v > 1U && bucket < 15) { v >>= 1U; ++bucket; }
            st.add(ST_HISTOGRAM + bucket);
        }
        st.add(ST_LOGICAL, static_cast<unsigned long long>(length));
    }
    if (committed) {
        const OrbitFastDev top = load_uniform(o.fast, r);
        const SN norm{top.zn_m, top.zn_e};
        s.derivative = derivative;
        s.iteration = iteration;
        s.ref = r;
        s.total = SC{top.z_re, top.z_im, top.z_e};
        s.total_norm = norm;
        s.margin = FE{top.margin_m, top.margin_e};
        s.previous_total_norm = norm;
        s.have_total = true;
    }
}

// The plain pass over a freshly set-up pixel (init_pixel): the plain-double
// block loop from the pixel's first loop top.  Returns true when the pixel is
// finished (interior at the cap: output and statistics written); otherwise
// `s` is a loop-top state to resume with run_linear, its derivative settled.
template <bool Planes, bool Stats>
FAV_HD bool plain_pixel(const RenderParams& p, size_t index, PixelState& s, SC dc, Counters<Stats>& st,
                        int32_t* dlog) {
    const int max_iter = p.max_iter;
    if (!p.fast_linear || !sc_normal(dc) || p.disable_bla || s.pixel_disable_bla) return false;
    DerivativeDebt debt;
    const PlainExit exit = run_plain_blocks<Planes, Stats>(p, s, dc, st, dlog, debt);
    if (exit == kPlainInterior) {  // the exact step to the cap, renderer.cpp:13516-13539
        s.iteration = max_iter;
        st.add(ST_LOGICAL);
        st.add(ST_EXACT);
    }
    if (s.iteration >= max_iter) {
        p.output[index] = encode_iteration(max_iter, p.output_bias);
        st.max(ST_MAX_PIXEL_ITER, static_cast<unsigned long long>(s.iteration));
        return true;
    }
    settle_derivative<Planes>(p, dlog, s, debt);
    return false;
}

// Fast path: the loop of renderer.cpp:12889-13536 restricted to iterations
// that end in a committed linear BLA block or in an exact perturbation step.
// Any other branch (glitch, escape, rebase, reference escape, cubic order,
// plain-double tail, replay, retry) is left to the general kernel: the pixel
// is handed on with its state at that loop top.  Interior pixels that reach
// the cap finish here.  With planes, `dlog` is the wave's derivative log of
// kDerivativeLog bytes, private to the wave (unused without planes).
template <bool Planes, bool Stats>
FAV_HD Outcome linear_loop(const RenderParams& p, size_t index, PixelState& s, SC dc, Counters<Stats>& st,
                           int32_t* dlog, DerivativeDebt& debt) {
    const OrbitDev& o = p.orbit;
    const int max_iter = p.max_iter;
    const int mode = p.escape_radius_mode;
    if (!p.fast_linear) return kHandoff;
    const bool normal_blocks = sc_normal(dc) && !p.disable_bla && !s.pixel_disable_bla;
    while (s.iteration < max_iter) {
        if (normal_blocks) {
            settle_derivative<Planes>(p, dlog, s, debt);
            run_normal_blocks<Planes, Stats>(p, s, dc, st, dlog, debt);
            if (s.iteration >= max_iter) break;
        } else if (sc_is_zero(dc) && !p.disable_bla && !s.pixel_disable_bla) {
            run_reference_blocks<Planes, Stats>(p, s, st);
            if (s.iteration >= max_iter) break;
        }
        if (s.ref < 0 || s.ref >= o.size) return kHandoff;
        if (!s.have_total) loop_top_values(o, s.ref, s.delta, mode, s.total, s.margin, s.total_norm);
        s.have_total = false;
        const FE glitch{o.glitch_m[s.ref], o.glitch_e[s.ref]};
        if (glitch.m != 0.0 && fe_less_nonneg(FE{s.total_norm.m, s.total_norm.e}, glitch)) return kHandoff;
        if (sc_outside_escape(s.total_norm, mode) || fe_compare(s.margin, fe_zero()) > 0) return kHandoff;
        const SN delta_norm = sc_norm_squared(s.delta);
        if (sc_compare_norm(s.total_norm, delta_norm) < 0) return kHandoff;
        if (o.map_end > 1 && s.ref >= o.map_end) return kHandoff;
        const bool bla = !p.disable_bla && !s.pixel_disable_bla;
        const int order = p.approximation_order >= 3 && delta_norm.e < -115
            ? (delta_norm.e < -160 ? 1 : 2) : p.approximation_order;
        // With delta = dc = 0 (the pixel sitting on the reference) every
        // term of the cubic block but A*delta + B*dc is an exact zero, and so
        // is every term of its derivative but A*dz + B (renderer.cpp:8686-
        // 8795), so the block is applied as a linear one; only its lookup
        // uses the cubic table.
        const bool on_reference = sc_is_zero(s.delta) && sc_is_zero(dc);
        if (bla && order > 1 && !on_reference) return kHandoff;  // cubic blocks
        int k = -1, length = 0;
        if (bla) {
            const int remaining = max_iter - s.iteration;
            const FE dn{delta_norm.m, delta_norm.e};
            if (order <= 1) {
                k = lookup_level(p.linear, p.linear_dir, s.ref, dn,
                                 p.max_linear_length < remaining ? p.max_linear_length : remaining);
                if (k >= 0) length = p.linear[k].length;
            } else {
                k = lookup_level(p.cubic, p.cubic_dir, s.ref, dn,
                                 p.max_bla_length < remaining ? p.max_bla_length : remaining);
                if (k >= 0) length = p.cubic[k].length;
            }
        }
        const bool block = length > 0 && s.ref + length <= max_iter;
        if (bla && !block && delta_norm.e > -90) return kHandoff;  // plain-double tail
        if (block) {
            SC a, b;
            if (order <= 1) {
                const LinearStepDev& step = p.linear[k];
                a = SC{step.a_re, step.a_im, step.a_e};
                b = SC{step.b_re, step.b_im, step.b_e};
            } else {
                a = coef(p.cubic[k], 0);
                b = coef(p.cubic[k], 1);
            }
            const SC delta = sc_add(sc_mul(a, s.delta), sc_mul(b, dc));
            settle_derivative<Planes>(p, dlog, s, debt);
            SC derivative = s.derivative;
            if constexpr (Planes) derivative = sc_add(sc_mul(a, s.derivative), b);
            const int ref = s.ref + length;
            SC endpoint;
            FE endpoint_margin;
            SN endpoint_norm;
            loop_top_values(o, ref, delta, mode, endpoint, endpoint_margin, endpoint_norm);
            if (!sc_finite(delta) || !sc_finite(endpoint)) return kHandoff;
            if (sc_compare_norm(endpoint_norm, SN{0.75, 2}) >= 0) return kHandoff;
            // commit, renderer.cpp:13499-13511
            s.derivative = derivative;
            s.delta = delta;
            s.ref = ref;
            s.iteration += length;
            s.total = endpoint;
            s.total_norm = endpoint_norm;
            s.margin = endpoint_margin;
            s.previous_total_norm = endpoint_norm;
            s.have_total = true;
            st.add(ST_BLA_BLOCKS);
            st.add(order <= 1 ? ST_LINEAR_BLOCKS : ST_CUBIC_BLOCKS);
            st.add(ST_SERIES_JUMPS);
            if constexpr (Stats) {
                int bucket = 0;
                unsigned v = static_cast<unsigned>(length);
                while (v > 1U && bucket < 15) { v >>= 1U; ++bucket; }
                st.add(ST_HISTOGRAM + bucket);
            }
            st.add(ST_LOGICAL, static_cast<unsigned long long>(length));
            continue;
        }
        // exact perturbation step, renderer.cpp:13516-13535; a step that
        // reaches the cap leaves an interior pixel, which reads no derivative
        const SC z = load_z(o, s.ref);
        if constexpr (Planes) {
            if (s.iteration + 1 < max_iter) {
                settle_derivative<Planes>(p, dlog, s, debt);
                s.derivative = sc_add(sc_double(sc_mul(s.total, s.derivative)), kOne);
            }
        }
        s.delta = sc_add(sc_double(sc_mul(z, s.delta)), sc_add(sc_mul(s.delta, s.delta), dc));
        ++s.ref;
        ++s.iteration;
        s.previous_total_norm = s.total_norm;
        st.add(ST_LOGICAL);
        st.add(ST_EXACT);
    }
    p.output[index] = encode_iteration(max_iter, p.output_bias);
    st.max(ST_MAX_PIXEL_ITER, static_cast<unsigned long long>(s.iteration > 0 ? s.iteration : 0));
    return kDone;
}

// linear_loop with the derivative settled for a pixel handed on (the state
// the general kernel resumes from carries it); an interior pixel ends with
// kDone and reads none.
template <bool Planes, bool Stats>
FAV_HD Outcome run_linear(const RenderParams& p, size_t index, PixelState& s, SC dc, Counters<Stats>& st,
                          int32_t* dlog) {
    DerivativeDebt debt;
    const Outcome out = linear_loop<Planes, Stats>(p, index, s, dc, st, dlog, debt);
    if (out == kHandoff) settle_derivative<Planes>(p, dlog, s, debt);
    return out;
}

// The general loop of render_bla_impl, renderer.cpp:12889-13544, resumed
// from a loop-top state.  With InlineTail the plain-double tail runs here;
// otherwise the pixel is handed to the tail kernel at the tail entry
// (renderer.cpp:13283-13290) and kTail is returned.
template <bool Planes, bool Stats, bool InlineTail>
FAV_HD Outcome run_general(const RenderParams& p, size_t index, PixelState& s, SC dc, Counters<Stats>& st) {
    const OrbitDev& o = p.orbit;
    const PlaneOut planes = plane_out(p);
    const int max_iter = p.max_iter;
    const int mode = p.escape_radius_mode;
    const FE spacing{p.pixel_spacing_m, p.pixel_spacing_e};
    bool escaped = false;
    SC tortoise = sc_zero();
    int cycle_power = 1, cycle_length = 0;
    bool cycle_ready = false;
    Budget budget{&p, s.ticks};
    bool deadline = false, unresolved = false;
    float out = 0.0f;
    const int cycle_min = p.strict_cycle ? 2048 : 512;
    const int cycle_margin = p.strict_cycle ? 78 : 60;
    // How the loop ended, when the end's work runs after it (once for all
    // the lanes of a wave that end at different iterations): the escape
    // (kEndEncoded: reference index out of range, the count as is;
    // kEndSmooth: the smooth value of s.total_norm), or the direct scaled
    // recurrence (kEndExactTail).
    enum { kEndNone, kEndEncoded, kEndSmooth, kEndExactTail } end = kEndNone;

    while (s.iteration < max_iter) {
        if (budget.expired()) {
            out = __builtin_nan(""); deadline = true; unresolved = true;
            st.add(ST_DEADLINE); st.add(ST_UNRESOLVED);
            break;
        }
        if (s.ref < 0 || s.ref >= o.size) {
            end = kEndEncoded;
            break;
        }
        if (!s.have_total) loop_top_values(o, s.ref, s.delta, mode, s.total, s.margin, s.total_norm);
        s.have_total = false;
        // glitch: |Z+d|^2 < 1e-7 |Z|^2, renderer.cpp:12931-12988 (the
        // threshold table holds 0 where |Z|^2 is zero or not finite)
        const FE glitch{o.glitch_m[s.ref], o.glitch_e[s.ref]};
        if (glitch.m != 0.0 && fe_less_nonneg(FE{s.total_norm.m, s.total_norm.e}, glitch)) {
            st.add(ST_GLITCH);
            if (p.allow_recovery && sc_finite(s.total) && ++s.rebases <= 64) {
                s.delta = s.total;
                s.ref = 0;
                continue;
            }
            out = __builtin_nan("");
            unresolved = true;
            st.add(ST_UNRESOLVED);
            break;
        }
        // escape, renderer.cpp:12989-13006
        if (sc_outside_escape(s.total_norm, mode) || fe_compare(s.margin, fe_zero()) > 0) {
            end = kEndSmooth;
            break;
        }
        // Brent cycle check, renderer.cpp:13016-13042
        if (p.cycle_enabled && s.iteration >= cycle_min && (s.iteration & 31) == 0
            && sc_compare_norm(s.total_norm, SN{0.75, 2}) < 0) {
            if (!cycle_ready) {
                tortoise = s.total; cycle_power = 1; cycle_length = 0; cycle_ready = true;
            } else {
                const SN distance = sc_norm_squared(sc_sub(s.total, tortoise));
                const int scale_exponent = s.total_norm.e > 0 ? s.total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale_exponent - cycle_margin) {
                    st.add(ST_CYCLE_INSIDE);
                    out = encode_iteration(max_iter, p.output_bias);
                    escaped = true;
                    break;
                }
                ++cycle_length;
                if (cycle_length >= cycle_power) {
                    tortoise = s.total;
                    cycle_power = cycle_power * 2 < (1 << 20) ? cycle_power * 2 : (1 << 20);
                    cycle_length = 0;
                }
            }
        }
        // rebase when |Z+d| < |d|, renderer.cpp:13047-13153
        const SN delta_norm = sc_norm_squared(s.delta);
        if (sc_compare_norm(s.total_norm, delta_norm) < 0) {
            if (++s.rebases > 64) {
                end = kEndExactTail;
                break;
            }
            s.delta = s.total;
            s.ref = 0;
            continue;
        }
        // the reference orbit escaped: direct recurrence, renderer.cpp:13155-13250
        if (o.map_end > 1 && s.ref >= o.map_end) {
            end = kEndExactTail;
            break;
        }
        // BLA selection, renderer.cpp:13252-13273
        const FE dn{delta_norm.m, delta_norm.e};
        const int remaining = max_iter - s.iteration;
        const int order = p.approximation_order >= 3 && delta_norm.e < -115
            ? (delta_norm.e < -160 ? 1 : 2) : p.approximation_order;
        int linear_k = -1, cubic_k = -1, map_length = 0;
        if (!p.disable_bla && !s.pixel_disable_bla) {
            if (order <= 1) {
                linear_k = lookup_level(p.linear, p.linear_dir, s.ref, dn,
                                        p.max_linear_length < remaining ? p.max_linear_length : remaining);
                if (linear_k >= 0) map_length = p.linear[linear_k].length;
            } else {
                cubic_k = lookup_level(p.cubic, p.cubic_dir, s.ref, dn,
                                       p.max_bla_length < remaining ? p.max_bla_length : remaining);
                if (cubic_k >= 0) map_length = p.cubic[cubic_k].length;
            }
        }
        // plain-double tail, renderer.cpp:13283-13362
        if (!p.disable_bla && !s.pixel_disable_bla && (map_length <= 0 || s.ref + map_length > max_iter)
            && delta_norm.e > -90) {
            st.add(ST_TAIL_PIXELS);
            if constexpr (!InlineTail) {
                s.ticks = budget.ticks;
                return kTail;  // the tail kernel continues from (delta, ref, iteration, total_norm)
            } else {
                const int tail_start = s.iteration;
                const bool pathological = double_tail<Planes, Stats>(
                    p, planes, index, dc, s.iteration, s.ref, s.delta, s.derivative, budget, deadline, unresolved,
                    st, s.total_norm, out);
                if (deadline || unresolved) break;
                const unsigned long long steps =
                    static_cast<unsigned long long>(s.iteration - tail_start > 0 ? s.iteration - tail_start : 0);
                st.add(ST_TAIL_STEPS, steps);
                st.max(ST_MAX_TAIL_STEPS, steps);
                if (pathological) {
                    s.delta = dc;
                    s.derivative = kOne;
                    s.ref = 1;
                    s.iteration = 1;
                    s.have_total = false;
                    s.pixel_disable_bla = true;
                    continue;
                }
                escaped = true;
                break;
            }
        }
        // BLA jump with endpoint guard and replay, renderer.cpp:13363-13514
        if (map_length > 0 && s.ref + map_length <= max_iter) {
            st.add(ST_BLA_BLOCKS);
            st.add(linear_k >= 0 ? ST_LINEAR_BLOCKS : ST_CUBIC_BLOCKS);
            st.add(ST_SERIES_JUMPS);
            if constexpr (Stats) {
                int bucket = 0;
                unsigned v = static_cast<unsigned>(map_length);
                while (v > 1U && bucket < 15) { v >>= 1U; ++bucket; }
                st.add(ST_HISTOGRAM + bucket);
            }
            const SC previous_delta = s.delta, previous_derivative = s.derivative;
            const int previous_ref = s.ref, previous_iteration = s.iteration;
            if (linear_k >= 0) {
                const LinearStepDev& step = p.linear[linear_k];
                const SC a{step.a_re, step.a_im, step.a_e}, b{step.b_re, step.b_im, step.b_e};
                s.delta = sc_add(sc_mul(a, previous_delta), sc_mul(b, dc));
                if constexpr (Planes) s.derivative = sc_add(sc_mul(a, previous_derivative), b);
            } else {
                const CubicStepDev& step = p.cubic[cubic_k];
                s.delta = apply_cubic(step, previous_delta, dc, order);
                if constexpr (Planes)
                    s.derivative = apply_cubic_derivative(step, previous_delta, dc, previous_derivative, order);
            }
            s.ref += map_length;
            s.iteration += map_length;
            SC endpoint;
            FE endpoint_margin;
            SN endpoint_norm;
            loop_top_values(o, s.ref, s.delta, mode, endpoint, endpoint_margin, endpoint_norm);
            if (!sc_finite(s.delta) || !sc_finite(endpoint)) {
                s.delta = previous_delta; s.derivative = previous_derivative;
                s.ref = previous_ref; s.iteration = previous_iteration;
                st.add(ST_BLA_RETRIES);
                if (!s.pixel_disable_bla) st.add(ST_BLA_DISABLED);
                s.pixel_disable_bla = true;
                s.have_total = false;
                continue;
            }
            if (sc_compare_norm(endpoint_norm, SN{0.75, 2}) >= 0) {
                s.delta = previous_delta; s.derivative = previous_derivative;
                s.ref = previous_ref; s.iteration = previous_iteration;
                SC replay_total = sc_zero();
                SN replay_norm{0.0, 0};
                FE replay_margin{0.0, 0};
                bool retry = false;
                for (int replay = 0; replay < map_length && s.iteration < max_iter; ++replay) {
                    if (budget.expired()) {
                        out = __builtin_nan(""); deadline = true; unresolved = true;
                        st.add(ST_DEADLINE); st.add(ST_UNRESOLVED);
                        break;
                    }
                    const SC zc = load_z(o, s.ref);
                    if constexpr (Planes)
                        s.derivative = sc_add(sc_double(sc_mul(sc_add(zc, s.delta), s.derivative)), kOne);
                    s.delta = sc_add(sc_double(sc_mul(zc, s.delta)), sc_add(sc_mul(s.delta, s.delta), dc));
                    ++s.ref;
                    ++s.iteration;
                    st.add(ST_REPLAY); st.add(ST_LOGICAL);
                    replay_total = sc_add(load_z(o, s.ref), s.delta);
                    replay_margin = margin_at(o, s.ref, s.delta);
                    replay_norm = sn_from_margin(replay_margin, mode);
                    if (!sc_finite(s.delta) || !sc_finite(replay_total)) { retry = true; break; }
                    if (sc_outside_escape(replay_norm, mode) || fe_compare(replay_margin, fe_zero()) > 0) {
                        // the escape's values, as the loop-top escape finds them
                        s.previous_total_norm = s.total_norm;
                        s.total = replay_total;
                        s.total_norm = replay_norm;
                        end = kEndSmooth;
                        break;
                    }
                }
                if (deadline || end != kEndNone) break;
                if (retry) {
                    s.delta = previous_delta; s.derivative = previous_derivative;
                    s.ref = previous_ref; s.iteration = previous_iteration;
                    st.add(ST_BLA_RETRIES);
                    if (!s.pixel_disable_bla) st.add(ST_BLA_DISABLED);
                    s.pixel_disable_bla = true;
                    s.have_total = false;
                    continue;
                }
                s.total = replay_total;
                s.total_norm = replay_norm;
                s.margin = replay_margin;
                s.previous_total_norm = s.total_norm;
                s.have_total = true;
            } else {
                s.total = endpoint;
                s.total_norm = endpoint_norm;
                s.margin = endpoint_margin;
                s.previous_total_norm = s.total_norm;
                s.have_total = true;
                st.add(ST_LOGICAL, static_cast<unsigned long long>(map_length));
            }
            continue;
        }
        // exact perturbation step, renderer.cpp:13516-13535
        const SC z = load_z(o, s.ref);
        const SN prior_norm = s.total_norm;
        if constexpr (Planes) s.derivative = sc_add(sc_double(sc_mul(sc_add(z, s.delta), s.derivative)), kOne);
        s.delta = sc_add(sc_double(sc_mul(z, s.delta)), sc_add(sc_mul(s.delta, s.delta), dc));
        ++s.ref;
        ++s.iteration;
        s.previous_total_norm = prior_norm;
        st.add(ST_LOGICAL); st.add(ST_EXACT);
    }
    if (end == kEndEncoded || end == kEndSmooth) {
        out = end == kEndSmooth ? smooth_escape(s.iteration, s.total_norm, p.output_bias)
                                : encode_iteration(s.iteration, p.output_bias);
        if constexpr (Planes)
            store_escape(planes, index, s.iteration, s.total, s.total_norm, s.previous_total_norm, s.derivative,
                         spacing);
        escaped = true;
    } else if (end == kEndExactTail) {
        exact_tail<Planes, Stats>(p, planes, index, dc, s.iteration, s.total, s.total_norm, s.derivative, budget,
                                  deadline, unresolved, st, out);
        if (!deadline) escaped = true;
    }
    if (!escaped && !deadline && !unresolved) out = encode_iteration(max_iter, p.output_bias);
    p.output[index] = out;
    st.max(ST_MAX_PIXEL_ITER, static_cast<unsigned long long>(s.iteration > 0 ? s.iteration : 0));
    return kDone;
}

// The tail kernel's part: render_scaled_double_tail from the handed-on state
// and the bookkeeping the caller does after it (renderer.cpp:13283-13362).
template <bool Planes, bool Stats>
FAV_HD void run_tail(const RenderParams& p, size_t index, PixelState& s, SC dc, Counters<Stats>& st) {
    Budget budget{&p, s.ticks};
    bool deadline = false, unresolved = false;
    float out = 0.0f;
    const int tail_start = s.iteration;
    (void)double_tail<Planes, Stats>(p, plane_out(p), index, dc, s.iteration, s.ref, s.delta, s.derivative, budget,
                                     deadline, unresolved, st, s.total_norm, out);
    // Both ways double_tail reports a pathological tail also mark the pixel
    // unresolved (renderer.cpp:12536-12544, 12644-12655), so the caller's
    // restart branch is never taken and the tail always ends the pixel.
    if (!deadline && !unresolved) {
        const unsigned long long steps =
            static_cast<unsigned long long>(s.iteration - tail_start > 0 ? s.iteration - tail_start : 0);
        st.add(ST_TAIL_STEPS, steps);
        st.max(ST_MAX_TAIL_STEPS, steps);
    }
    p.output[index] = out;
    st.max(ST_MAX_PIXEL_ITER, static_cast<unsigned long long>(s.iteration > 0 ? s.iteration : 0));
}

}  // namespace core
}  // namespace fav
