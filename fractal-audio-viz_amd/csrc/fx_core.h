// fx_core.h -- the FP32 "floatexp" perturbation path: pixel deltas carried
// as float mantissa pairs with one int32 exponent, renormalised after every
// operation.  The reference orbit, the BLA coefficients and validity radii
// stay in their FP64 tables (converted per use), and the escape, glitch and
// rebase decisions and the final log-log smoothing are taken in FP64 on the
// norms.  Unlike the FP64 path this is not the reference's arithmetic, so a
// small fraction of boundary pixels may escape at a different iteration; the
// caller measures that fraction against the FP64 kernel.
//
// Mandelbrot only.  Linear BLA blocks at every delta size (the FP64 path
// switches to the cubic table for large deltas), rebasing to the start of
// the orbit when |Z+delta| < |delta| or the reference escapes.
#pragma once

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
    int32_t* counts;  // escape iteration, max_iter for interior, -1 for an unresolved (NaN) pixel
};

// One pixel of the floatexp path.  `p` carries the FP64 path's tables and
// launch parameters (orbit, linear BLA tier, series, offsets).
FAV_HD void pixel(const RenderParams& p, const FxOut& out, int px, int py) {
    const OrbitDev& o = p.orbit;
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    const int max_iter = p.max_iter;
    const int mode = p.escape_radius_mode;
    if (o.size < 2) {  // no orbit to perturb against
        out.field[index] = __builtin_nanf("");
        out.counts[index] = -1;
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
    while (iteration < max_iter) {
        if (!p.disable_bla) {
            run_blocks(p, delta, dc, ref, iteration, last);
            if (iteration >= max_iter) break;
        }
        if (ref >= last) {  // continue against the orbit's start, Z_0 = 0
            delta = add(from_sc(core::load_z(o, ref)), delta);
            ref = 0;
        }
        const FX z = from_sc(core::load_z(o, ref));
        const FX total = add(z, delta);
        const SN norm = norm2(total);
        if (sc_outside_escape(norm, mode)) {
            out.field[index] = core::smooth_escape(iteration, norm, p.output_bias);
            out.counts[index] = iteration;
            return;
        }
        const FE glitch{o.glitch_m[ref], o.glitch_e[ref]};
        if (glitch.m != 0.0 && fe_less_nonneg(FE{norm.m, norm.e}, glitch)) {
            if (!p.allow_recovery || ++rebases > 64) {
                out.field[index] = __builtin_nanf("");
                out.counts[index] = -1;
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
        const int k = p.disable_bla ? -1
            : core::lookup_level(p.linear, p.linear_dir, ref, FE{dn.m, dn.e},
                                 p.max_linear_length < remaining ? p.max_linear_length : remaining);
        if (k >= 0 && ref + p.linear[k].length <= max_iter && ref + p.linear[k].length <= last) {
            const LinearStepDev& step = p.linear[k];
            const FX a = from_sc(SC{step.a_re, step.a_im, step.a_e});
            const FX b = from_sc(SC{step.b_re, step.b_im, step.b_e});
            delta = add(mul(a, delta), mul(b, dc));
            ref += step.length;
            iteration += step.length;
            continue;
        }
        // exact step d' = 2 Z d + d^2 + dc
        delta = add(twice(mul(z, delta)), add(mul(delta, delta), dc));
        ++ref;
        ++iteration;
    }
    out.field[index] = core::encode_iteration(max_iter, p.output_bias);
    out.counts[index] = max_iter;
}

}  // namespace fx
}  // namespace fav
