// fx_core.h -- the FP32 "floatexp" perturbation path: pixel deltas carried
// as float mantissa pairs with one int32 exponent, renormalised after every
// operation.  The reference orbit, the BLA coefficients and validity radii
// stay in their FP64 tables (converted per use), and the escape, glitch and
// rebase decisions and the final log-log smoothing are taken in FP64 on the
// norms.  Unlike the FP64 path this is not the reference's arithmetic, so a
// small fraction of boundary pixels may escape at a different iteration; the
// caller measures that fraction against the FP64 kernel.
//
// Mandelbrot map: linear BLA blocks at every delta size (the FP64 path
// switches to the cubic table for large deltas), rebasing to the start of
// the orbit when |Z+delta| < |delta| or the reference escapes, and the FP64
// plain-double tail where the FP64 path enters it.  Julia, burning ship and
// tricorn (alt_pixel): the exact perturbation step of each map (the burning
// ship's sign branches with the axis-crossing reconstruction) and the real
// 2x2 linear BLA blocks of the FP64 alternate tables, in the same mantissa
// arithmetic.
#pragma once

#include "alt_core.h"
#include "field_core.h"

namespace fav {
namespace fx {

struct FX {
    float re, im;
    int e;
};

FAV_HD int frexp_exp_f(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_frexp_expf(v);
#else
    int e = 0;
    (void)std::frexp(v, &e);
    return e;
#endif
}

FAV_HD float ldexp_f(float v, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ldexpf(v, e);
#else
    return std::ldexp(v, e);
#endif
}

// Larger part scaled into [0.5, 1); zero maps to exponent 0.
FAV_HD FX normalize(float re, float im, int e) {
    // the larger magnitude (operands are finite): one max instruction
    const float magnitude = __builtin_fmaxf(__builtin_fabsf(re), __builtin_fabsf(im));
    const int shift = magnitude == 0.0F ? 0 : frexp_exp_f(magnitude);
    return FX{ldexp_f(re, -shift), ldexp_f(im, -shift), magnitude == 0.0F ? 0 : e + shift};
}

FAV_HD FX from_sc(SC v) { return normalize(static_cast<float>(v.re), static_cast<float>(v.im), v.e); }
FAV_HD bool is_zero(FX v) { return v.re == 0.0F && v.im == 0.0F; }

FAV_HD FX mul(FX a, FX b) {
    return normalize(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re, a.e + b.e);
}

FAV_HD FX add(FX a, FX b) {
    if (is_zero(a)) return b;
    if (is_zero(b)) return a;
    const bool swap = b.e > a.e;
    const FX larger = swap ? b : a;
    const FX smaller = swap ? a : b;
    const int d = larger.e - smaller.e;
    if (d > 30) return larger;  // below float resolution of the larger value
    return normalize(larger.re + ldexp_f(smaller.re, -d), larger.im + ldexp_f(smaller.im, -d), larger.e);
}

FAV_HD FX twice(FX v) { return is_zero(v) ? v : FX{v.re, v.im, v.e + 1}; }

// |v|^2 as an FP64 ScaledNorm (mantissa in [0.5, 1)).
FAV_HD SN norm2(FX v) {
    if (is_zero(v)) return SN{0.0, 0};
    const double n = static_cast<double>(v.re) * v.re + static_cast<double>(v.im) * v.im;
    const int shift = frexp_exp_nz(n);
    return SN{ldexp_(n, -shift), 2 * v.e + shift};
}

// Linear blocks in the negligible-perturbation regime, decided for the
// whole wave as in the FP64 block loop (field_core.h run_normal_blocks):
// the loop top's escape / glitch / endpoint outcomes are the orbit's flags,
// all lanes share the reference index and BLA level, and the table reads are
// scalar.  Returns when any lane needs the general loop below.
FAV_HD void run_blocks(const RenderParams& p, FX& delta, FX dc, int& ref, int& iteration, int last) {
    const OrbitDev& o = p.orbit;
    const LevelDir& dir = p.linear_dir;
    const int max_iter = p.max_iter;
    int r = core::wave_first(ref);
    if (!core::wave_all(ref == r && iteration < max_iter && !is_zero(delta)) || r <= 0 || r >= last) return;
    OrbitFastDev top = core::load_uniform(o.fast, r);
    FX d = delta;
    int it = iteration;
    bool committed = false;
    for (;;) {
        if ((top.flags & (kOrbitGlitchFlag | kOrbitEscapeFlag)) != 0) break;
        const SN dn = norm2(d);
        const int remaining = max_iter - it;
        const int limit = p.max_linear_length < remaining ? p.max_linear_length : remaining;
        const int lim = core::wave_first(limit);
        const bool ok = d.e <= top.small_exp && sc_compare_norm(SN{top.zn_m, top.zn_e}, dn) >= 0 && limit == lim;
        if (!core::wave_all(ok)) break;
        if (lim <= 0 || dir.count == 0 || r > dir.size[0]) break;
        int h = core::highest_level(lim);
        if (h > dir.count - 1) h = dir.count - 1;
        const int offset = r - 1;
        if (offset != 0 && __builtin_ctz(static_cast<unsigned>(offset)) < h)
            h = __builtin_ctz(static_cast<unsigned>(offset));
        int k = -1;
        bool mixed = false;
        for (int level = h; level >= 0; --level) {
            const int index = offset >> level;
            if (index >= dir.size[level]) continue;
            const int64_t kk = dir.offset[level] + index;
            const LinearStepDev cand = core::load_uniform(p.linear, kk);
            if ((cand.pad[0] & kStepUsable) == 0) continue;
            const bool fits = dn.e != cand.r2_e ? dn.e < cand.r2_e : dn.m < cand.r2_m;
            if (core::wave_all(fits)) {
                k = static_cast<int>(kk);
                break;
            }
            if (core::wave_any(fits)) {
                mixed = true;
                break;
            }
        }
        if (k < 0 || mixed) break;
        const LinearStepDev step = core::load_uniform(p.linear, k);
        const int end = r + step.length;
        if (end > max_iter || end > last) break;
        const OrbitFastDev next = core::load_uniform(o.fast, end);
        if ((next.flags & kOrbitEndpointFlag) != 0) break;
        const FX a = from_sc(SC{step.a_re, step.a_im, step.a_e});
        const FX b = from_sc(SC{step.b_re, step.b_im, step.b_e});
        const FX d1 = add(mul(a, d), mul(b, dc));
        if (!core::wave_all(!is_zero(d1))) break;
        d = d1;
        it += step.length;
        r = end;
        top = next;
        committed = true;
        if (!core::wave_all(it < max_iter) || r >= last) break;
    }
    if (committed) {
        delta = d;
        iteration = it;
        ref = r;
    }
}

struct FxOut {
    float* field;
    int32_t* counts;  // escape iteration, max_iter for interior, -1 for an unresolved (NaN) pixel; may be null
    // 1: where the FP64 path enters its plain-double tail (no BLA block and
    // |d|^2 > 2^-90), the pixel continues in that tail in FP64
    int32_t double_tail;
};

// The FP64 plain-double tail of field_core.h from the floatexp state at a
// loop top (delta converted exactly to a ScaledComplex, `norm` = |Z+d|^2).
// Returns false when the tail gave up (the FP64 path's restart with BLA
// off); otherwise the pixel is finished: field and count written.
FAV_HD bool finish_in_double_tail(const RenderParams& p, const FxOut& out, size_t index, SC dc, FX delta, int& ref,
                                  int& iteration, SN norm) {
    core::Counters<false> st;
    core::Budget budget{&p, 0U};
    bool deadline = false, unresolved = false;
    SC derivative{1.0, 0.0, 0};
    float value = 0.0F;
    const SC d{static_cast<double>(delta.re), static_cast<double>(delta.im), delta.e};
    const bool gave_up = core::double_tail<false, false>(p, core::plane_out(p), index, dc, iteration, ref, d,
                                                         derivative, budget, deadline, unresolved, st, norm, value);
    if (gave_up && !deadline && !unresolved) return false;
    const float interior = core::encode_iteration(p.max_iter, p.output_bias);
    out.field[index] = value;
    if (out.counts) out.counts[index] = value != value ? -1 : (value == interior || iteration >= p.max_iter ? p.max_iter : iteration);
    return true;
}

// One pixel of the floatexp path.  `p` carries the FP64 path's tables and
// launch parameters (orbit, linear BLA tier, series, offsets).
FAV_HD void pixel(const RenderParams& p, const FxOut& out, int px, int py) {
    const OrbitDev& o = p.orbit;
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    const int max_iter = p.max_iter;
    const int mode = p.escape_radius_mode;
    if (o.size < 2) {  // no orbit to perturb against
        out.field[index] = __builtin_nanf("");
        if (out.counts) out.counts[index] = -1;
        return;
    }
    const SC dc_sc = core::pixel_dc(p, px, py);
    const FX dc = from_sc(dc_sc);
    FX delta = dc;
    int ref = 1, iteration = 1, rebases = 0;
    if (p.series.available
        && sc_compare_norm(sc_norm_squared(dc_sc), SN{p.series.radius2_m, p.series.radius2_e}) <= 0) {
        const int n = p.series.count;  // image-wide series evaluated in FP64
        SC value{p.series.re[n - 1], p.series.im[n - 1], p.series.e[n - 1]};
        for (int k = n - 1; k > 1; --k)
            value = sc_add(sc_mul(value, dc_sc), SC{p.series.re[k - 1], p.series.im[k - 1], p.series.e[k - 1]});
        delta = from_sc(sc_mul(value, dc_sc));
        ref = p.series.iteration;
        iteration = p.series.iteration;
    }
    // last usable reference index: the reference's escape (map_end) or the
    // end of the stored orbit
    const int last = o.map_end > 0 && o.map_end < o.size ? o.map_end : o.size - 1;
    bool bla_off = p.disable_bla != 0;  // also after a double tail gave up
    while (iteration < max_iter) {
        if (!bla_off) {
            run_blocks(p, delta, dc, ref, iteration, last);
            if (iteration >= max_iter) break;
        }
        if (ref >= last) {
            // the reference escaped: the direct recurrence on the full value in
            // FP64, as the FP64 path continues (renderer.cpp:13155-13250)
            const SC full = sc_add(core::load_z(o, ref), SC{delta.re, delta.im, delta.e});
            core::Counters<false> st;
            core::Budget budget{&p, 0U};
            bool deadline = false, unresolved = false;
            SC derivative{1.0, 0.0, 0};
            float value = 0.0F;
            core::exact_tail<false, false>(p, core::plane_out(p), index, dc_sc, iteration, full,
                                           sc_norm_squared(full), derivative, budget, deadline, unresolved, st,
                                           value);
            const float interior = core::encode_iteration(max_iter, p.output_bias);
            out.field[index] = value;
            if (out.counts) out.counts[index] = value != value ? -1 : (value == interior || iteration >= max_iter ? max_iter : iteration);
            return;
        }
        const FX z = from_sc(core::load_z(o, ref));
        const FX total = add(z, delta);
        const SN norm = norm2(total);
        if (sc_outside_escape(norm, mode)) {
            out.field[index] = core::smooth_escape(iteration, norm, p.output_bias);
            if (out.counts) out.counts[index] = iteration;
            return;
        }
        const FE glitch{o.glitch_m[ref], o.glitch_e[ref]};
        if (glitch.m != 0.0 && fe_less_nonneg(FE{norm.m, norm.e}, glitch)) {
            if (!p.allow_recovery || ++rebases > 64) {
                out.field[index] = __builtin_nanf("");
                if (out.counts) out.counts[index] = -1;
                return;
            }
            delta = total;
            ref = 0;
            continue;
        }
        const SN dn = norm2(delta);
        if (ref > 0 && sc_compare_norm(norm, dn) < 0) {  // |Z+d| < |d|: rebase (FP64 test)
            delta = total;
            ref = 0;
            continue;
        }
        const int remaining = max_iter - iteration;
        const int k = bla_off ? -1
            : core::lookup_level(p.linear, p.linear_dir, ref, FE{dn.m, dn.e},
                                 p.max_linear_length < remaining ? p.max_linear_length : remaining);
        const bool block = k >= 0 && ref + p.linear[k].length <= max_iter && ref + p.linear[k].length <= last;
        if (block) {
            const LinearStepDev& step = p.linear[k];
            const FX a = from_sc(SC{step.a_re, step.a_im, step.a_e});
            const FX b = from_sc(SC{step.b_re, step.b_im, step.b_e});
            delta = add(mul(a, delta), mul(b, dc));
            ref += step.length;
            iteration += step.length;
            continue;
        }
        // no block: where the FP64 path starts its plain-double tail
        // (renderer.cpp:13283-13290), the pixel finishes in that tail in FP64
        if (out.double_tail && !bla_off && dn.e > -90) {
            if (finish_in_double_tail(p, out, index, dc_sc, delta, ref, iteration, norm)) return;
            delta = dc;  // the tail gave up: restart without BLA, as the FP64 path does
            ref = 1;
            iteration = 1;
            bla_off = true;
            continue;
        }
        // exact step d' = 2 Z d + d^2 + dc
        delta = add(twice(mul(z, delta)), add(mul(delta, delta), dc));
        ++ref;
        ++iteration;
    }
    out.field[index] = core::encode_iteration(max_iter, p.output_bias);
    if (out.counts) out.counts[index] = max_iter;
}

// ---- alternate formulas ------------------------------------------------------

// A real value as a float mantissa and an int32 exponent (one part of an FX).
struct FS {
    float m;
    int e;
};

FAV_HD FS fs_norm(float m, int e) {
    if (m == 0.0F) return FS{0.0F, 0};
    const int shift = frexp_exp_f(m);
    return FS{ldexp_f(m, -shift), e + shift};
}

FAV_HD FS fs_from_fe(FE v) { return fs_norm(static_cast<float>(v.m), v.e); }
FAV_HD FS fs_mul(FS a, FS b) { return fs_norm(a.m * b.m, a.e + b.e); }

FAV_HD FS fs_add(FS a, FS b) {
    if (a.m == 0.0F) return b;
    if (b.m == 0.0F) return a;
    const bool swap = b.e > a.e;
    const FS larger = swap ? b : a;
    const FS smaller = swap ? a : b;
    const int d = larger.e - smaller.e;
    if (d > 30) return larger;
    return fs_norm(larger.m + ldexp_f(smaller.m, -d), larger.e);
}

// The complex value (re, im), its parts aligned to the larger exponent.
FAV_HD FX from_parts(FS re, FS im) {
    if (re.m == 0.0F && im.m == 0.0F) return FX{0.0F, 0.0F, 0};
    const int e = re.m == 0.0F ? im.e : (im.m == 0.0F ? re.e : (re.e > im.e ? re.e : im.e));
    const float r = re.m == 0.0F ? 0.0F : (e - re.e > 30 ? 0.0F : ldexp_f(re.m, re.e - e));
    const float i = im.m == 0.0F ? 0.0F : (e - im.e > 30 ? 0.0F : ldexp_f(im.m, im.e - e));
    return normalize(r, i, e);
}

FAV_HD FX conj(FX v) { return FX{v.re, -v.im, v.e}; }
FAV_HD FX abs_parts(FX v) { return FX{__builtin_fabsf(v.re), __builtin_fabsf(v.im), v.e}; }

// apply_alternate_linear_bla (renderer.cpp:11399-11416) in mantissa
// arithmetic: d' = M d + P q with the eight FP64 coefficients rounded to
// float mantissas.
FAV_HD FX alt_apply(const AltStepDev& s, FX d, FX q) {
    const FS dr{d.re, d.e}, di{d.im, d.e}, qr{q.re, q.e}, qi{q.im, q.e};
    FS c[8];
    for (int i = 0; i < 8; ++i) c[i] = fs_from_fe(FE{s.m[i], s.e[i]});
    const FS rr = fs_add(fs_add(fs_mul(c[0], dr), fs_mul(c[1], di)), fs_add(fs_mul(c[4], qr), fs_mul(c[5], qi)));
    const FS ri = fs_add(fs_add(fs_mul(c[2], dr), fs_mul(c[3], di)), fs_add(fs_mul(c[6], qr), fs_mul(c[7], qi)));
    return from_parts(rr, ri);
}

// One exact perturbation step of the alternate map at reference index i
// (the steps of alternate_delta_step, renderer.cpp:11933-12034, and of the
// reference's mixed-precision device kernel, renderer.cpp:1537-1701).
FAV_HD FX alt_step(const AltRenderParams& p, int i, FX d, FX pd) {
    const SC z_sc = alt::load_z(p, i);
    const FX z = from_sc(z_sc);
    if (p.formula == 3) {
        const FX cz = conj(z), cd = conj(d);
        return add(twice(mul(cz, cd)), add(mul(cd, cd), pd));
    }
    if (p.formula == 2) {
        // the sign tests per component on FloatExp sums, as the FP64 path
        // takes them: the shared exponent of a float pair cannot hold a part
        // of Z far below the other (Im Z near 0 on the real axis) against a
        // delta part above it
        const FEC zp = sc_parts(z_sc);
        const FE tr = fe_add(zp.re, fe_from_parts(static_cast<double>(d.re), d.e));
        const FE ti = fe_add(zp.im, fe_from_parts(static_cast<double>(d.im), d.e));
        const bool crosses_real = (fe_compare(zp.re, fe_zero()) >= 0) != (fe_compare(tr, fe_zero()) >= 0);
        const bool crosses_imag = (fe_compare(zp.im, fe_zero()) >= 0) != (fe_compare(ti, fe_zero()) >= 0);
        if (crosses_real || crosses_imag) {
            // the crossing perturbation is rebuilt from |Z+d|^2 minus the next
            // reference value: a difference of values of the reference's
            // size, taken in FP64 as the FP64 path takes it (a float mantissa
            // would round away the perturbation)
            return from_sc(alt::delta_step(p, i, SC{d.re, d.im, d.e}, SC{pd.re, pd.im, pd.e}));
        }
        const FX sd{zp.re.m < 0.0 ? -d.re : d.re, zp.im.m < 0.0 ? -d.im : d.im, d.e};
        return add(twice(mul(abs_parts(z), sd)), add(mul(sd, sd), pd));
    }
    return add(twice(mul(z, d)), add(mul(d, d), pd));
}

// |d|^2 above 2^kDoubleDelta hands an alternate-formula pixel to the FP64
// loop: the bound at which the Mandelbrot path's FP64 plain-double tail
// starts (renderer.cpp:13283-13290).
constexpr int kDoubleDelta = -90;

FAV_HD void alt_finish(const FxOut& out, size_t index, float value, int count) {
    out.field[index] = value;
    if (out.counts) out.counts[index] = count;
}

// One pixel of an alternate-formula field in floatexp: the control flow of
// render_alternate_reference_impl (alt_core.h pixel_body) without the cycle
// test and time budget; escape and block-validity decisions on FP64 norms of
// the floatexp totals.
FAV_HD void alt_pixel(const AltRenderParams& p, const FxOut& out, int px, int py) {
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    const bool julia = p.formula == 1;
    const int mode = p.escape_radius_mode;
    const double bias = p.output_bias;
    const int max_iter = p.max_iter;
    const SC dc_sc = alt::pixel_dc(p, px, py);
    const SC pd_sc = julia ? sc_zero() : dc_sc;
    const SN pnorm = sc_norm_squared(pd_sc);
    const FE pn{pnorm.m, pnorm.e};
    const FX pd = from_sc(pd_sc);
    FX delta = from_sc(dc_sc);
    int ref = julia ? 0 : 1, iteration = julia ? 0 : 1;
    if (p.series_available && sc_compare_norm(sc_norm_squared(dc_sc), SN{p.series.radius2_m, p.series.radius2_e}) <= 0) {
        delta = from_sc(alt::series_value(p.series, dc_sc));
        ref = p.series.iteration;
        iteration = p.series.iteration;
    }
    const SN start_norm = norm2(add(from_sc(alt::load_z(p, ref)), delta));
    if (sc_outside_escape(start_norm, mode)) {
        alt_finish(out, index, core::smooth_escape(iteration, start_norm, bias), iteration);
        return;
    }
    const SN limit{0.75, 2};
    while (iteration < max_iter) {
        const int next = ref + 1;
        const SN dn = norm2(delta);
        if (next >= p.orbit_size || dn.e > kDoubleDelta) {
            // past the reference's escape, or |d|^2 > 2^-90: the pixel
            // continues in the FP64 loop from this loop top (a float mantissa
            // cannot hold Z + d once d is no longer small against Z, nor the
            // pixel's offset from the parameter on the full value)
            const SC d{static_cast<double>(delta.re), static_cast<double>(delta.im), delta.e};
            int work = 0;
            const alt::Result r = alt::run_loop(p, dc_sc, pd_sc, pn, d, ref, iteration, sc_add(alt::load_z(p, ref), d),
                                                sn_from_margin(alt::margin(p, ref, d), mode), work);
            alt_finish(out, index, r.value, r.count);
            return;
        }
        const int remaining = max_iter - iteration;
        const int k = p.bla_enabled
            ? alt::lookup(p, ref, FE{dn.m, dn.e}, pn, p.max_block < remaining ? p.max_block : remaining) : -1;
        if (k >= 0 && ref + p.steps[k].length < p.orbit_size) {
            const AltStepDev& step = p.steps[k];
            const int end = ref + step.length;
            const FX cand = alt_apply(step, delta, pd);
            const FX ct = add(from_sc(alt::load_z(p, end)), cand);
            const SN cn = norm2(ct);
            if (!sc_outside_escape(cn, mode) && sc_compare_norm(cn, limit) < 0) {
                delta = cand;
                ref = end;
                iteration += step.length;
                continue;
            }
            // replay the block one exact step at a time
            for (int replay = 0; replay < step.length && iteration < max_iter; ++replay) {
                delta = alt_step(p, ref, delta, pd);
                ++ref;
                ++iteration;
                const SN n = norm2(add(from_sc(alt::load_z(p, ref)), delta));
                if (sc_outside_escape(n, mode)) {
                    alt_finish(out, index, core::smooth_escape(iteration, n, bias), iteration);
                    return;
                }
            }
            continue;
        }
        delta = alt_step(p, ref, delta, pd);
        ref = next;
        ++iteration;
        const SN n = norm2(add(from_sc(alt::load_z(p, ref)), delta));
        if (sc_outside_escape(n, mode)) {
            alt_finish(out, index, core::smooth_escape(iteration, n, bias), iteration);
            return;
        }
    }
    alt_finish(out, index, core::encode_iteration(max_iter, bias), max_iter);
}

}  // namespace fx
}  // namespace fav
