// alt_core.h -- per-pixel perturbation loop of the alternate deep formulas
// (render_alternate_reference_impl without planes, renderer.cpp:12064-12448):
// Julia (1), burning ship (2, sign-branch perturbation with exact
// axis-crossing reconstruction) and tricorn (3, conjugate map), with the
// real 2x2 linear BLA, exact replay of rejected blocks, continuation past
// the reference's escape and the Brent-style cycle test of non-strict renders.
//
// Written once for the gfx950 kernel (alt_kernel.hip) and compiled unchanged
// for the test-only CPU harness.  Same FloatExp / ScaledComplex operations
// in the same order as the reference, so escape counts are bit-identical.
#pragma once

#include "field_core.h"

namespace fav {
namespace alt {

FAV_HD FE fe_abs(FE v) { return FE{fabs_(v.m), v.e}; }
FAV_HD SC load_z(const AltRenderParams& p, int i) { return SC{p.re[i], p.im[i], p.e[i]}; }

// alternate_escape_margin / sc_escape_margin_with_reference_margin,
// renderer.cpp:8435-8451, 12036-12053: the MPFR margin plus 2 Re(Z conj d)
// plus |d|^2 as a ScaledNorm.
FAV_HD FE margin(const AltRenderParams& p, int i, SC delta) {
    const SC z = load_z(p, i);
    const FEC zp = sc_parts(z), dp = sc_parts(delta);
    const FE cross = fe_muld(fe_add(fe_mul(zp.re, dp.re), fe_mul(zp.im, dp.im)), 2.0);
    const SN dn = sc_norm_squared(delta);
    return fe_add(fe_add(FE{p.margin_m[i], p.margin_e[i]}, cross), FE{dn.m, dn.e});
}

// scaled_formula_step, renderer.cpp:11917-11931.
FAV_HD SC formula_step(int formula, SC v, SC parameter) {
    if (formula == 2) {
        const SC a{fabs_(v.re), fabs_(v.im), v.e};
        return sc_add(sc_mul(a, a), parameter);
    }
    const SC sq = formula == 3 ? sc_mul(sc_conj(v), sc_conj(v)) : sc_mul(v, v);
    return sc_add(sq, parameter);
}

// alternate_delta_step, renderer.cpp:11933-12034.
FAV_HD SC delta_step(const AltRenderParams& p, int i, SC delta, SC pdelta) {
    const SC z = load_z(p, i);
    if (p.formula == 3) {
        const SC cz = sc_conj(z), cd = sc_conj(delta);
        return sc_add(sc_double(sc_mul(cz, cd)), sc_add(sc_mul(cd, cd), pdelta));
    }
    if (p.formula == 2) {
        const FEC zp = sc_parts(z), dp = sc_parts(delta), pp = sc_parts(pdelta);
        const FE azr = fe_abs(zp.re), azi = fe_abs(zp.im);
        const FE tr = fe_add(zp.re, dp.re), ti = fe_add(zp.im, dp.im);
        const bool crosses_real = (fe_compare(zp.re, fe_zero()) >= 0) != (fe_compare(tr, fe_zero()) >= 0);
        const bool crosses_imag = (fe_compare(zp.im, fe_zero()) >= 0) != (fe_compare(ti, fe_zero()) >= 0);
        if (crosses_real || crosses_imag) {
            // rebuild |Z+d| from FloatExp parts and subtract the next
            // reference value, keeping the crossing perturbation
            const FE atr = fe_abs(tr), ati = fe_abs(ti);
            const FEC next = sc_parts(load_z(p, i + 1));
            const FEC param = sc_parts(SC{p.parameter_re, p.parameter_im, p.parameter_e});
            const FE por = fe_sub(param.re, next.re), poi = fe_sub(param.im, next.im);
            const FE nr = fe_add(fe_sub(fe_sqr(atr), fe_sqr(ati)), fe_add(por, pp.re));
            const FE ni = fe_add(fe_muld(fe_mul(atr, ati), 2.0), fe_add(poi, pp.im));
            return sc_from_fe(nr, ni);
        }
        const double rs = zp.re.m < 0.0 ? -1.0 : 1.0, is = zp.im.m < 0.0 ? -1.0 : 1.0;
        const FE rl = fe_sub(fe_muld(fe_mul(azr, dp.re), 2.0 * rs), fe_muld(fe_mul(azi, dp.im), 2.0 * is));
        const FE il = fe_add(fe_muld(fe_mul(azr, dp.im), 2.0 * is), fe_muld(fe_mul(azi, dp.re), 2.0 * rs));
        const FE sr = fe_sub(fe_sqr(dp.re), fe_sqr(dp.im));
        const FE si = fe_muld(fe_mul(dp.re, dp.im), 2.0 * rs * is);
        return sc_from_fe(fe_add(fe_add(rl, sr), pp.re), fe_add(fe_add(il, si), pp.im));
    }
    return sc_add(sc_double(sc_mul(z, delta)), sc_add(sc_mul(delta, delta), pdelta));
}

// AlternateLinearBlaLevels::lookup, renderer.cpp:8533-8565.  Returns the
// flat step index or -1.
FAV_HD int lookup(const AltRenderParams& p, int start, FE dn, FE pn, int max_length) {
    const LevelDir& d = p.dir;
    if (d.count == 0 || start < p.alt_start || max_length <= 0 || !fe_finite(dn) || !fe_finite(pn)) return -1;
    const int offset = start - p.alt_start;
    if (offset < 0 || offset >= d.size[0]) return -1;
    int h = core::highest_level(max_length);
    if (h > d.count - 1) h = d.count - 1;
    const FE input_r2{p.input_r2_m, p.input_r2_e};
    for (int level = h; level >= 0; --level) {
        if ((offset & ((1 << level) - 1)) != 0) continue;
        const int index = offset >> level;
        if (index >= d.size[level]) continue;
        const int64_t k = d.offset[level] + index;
        const AltStepDev& s = p.steps[k];
        const FE r2{s.r2_m, s.r2_e};
        if (s.length > 1 && fe_finite(r2) && fe_compare(dn, r2) < 0 && fe_compare(pn, input_r2) <= 0)
            return static_cast<int>(k);
    }
    return -1;
}

// apply_alternate_linear_bla, renderer.cpp:11399-11416.
FAV_HD SC apply(const AltStepDev& s, SC delta, SC parameter) {
    const FEC d = sc_parts(delta), q = sc_parts(parameter);
    const FE c0{s.m[0], s.e[0]}, c1{s.m[1], s.e[1]}, c2{s.m[2], s.e[2]}, c3{s.m[3], s.e[3]};
    const FE c4{s.m[4], s.e[4]}, c5{s.m[5], s.e[5]}, c6{s.m[6], s.e[6]}, c7{s.m[7], s.e[7]};
    const FE rr = fe_add(fe_add(fe_mul(c0, d.re), fe_mul(c1, d.im)), fe_add(fe_mul(c4, q.re), fe_mul(c5, q.im)));
    const FE ri = fe_add(fe_add(fe_mul(c2, d.re), fe_mul(c3, d.im)), fe_add(fe_mul(c6, q.re), fe_mul(c7, q.im)));
    return sc_from_fe(rr, ri);
}

// evaluate_image_series (value only), renderer.cpp:8611-8639.
FAV_HD SC series_value(const SeriesDev& s, SC dc) {
    const int n = s.count;
    if (n <= 1) return sc_zero();
    SC value{s.re[n - 1], s.im[n - 1], s.e[n - 1]};
    for (int k = n - 1; k > 1; --k) value = sc_add(sc_mul(value, dc), SC{s.re[k - 1], s.im[k - 1], s.e[k - 1]});
    return sc_mul(value, dc);
}

struct AltBudget {  // render_time_budget_expired, renderer.cpp:11878-11890
    const AltRenderParams* p;
    unsigned ticks;
    FAV_HD bool expired() {
        if (p->budget_ticks == 0) return false;
        if ((ticks++ & 255U) != 0U) return false;
        if (core::flag_load(p->budget_flag)) return true;
        if (core::clock_ticks() - *p->budget_start >= p->budget_ticks) {
            core::flag_store(p->budget_flag);
            return true;
        }
        return false;
    }
};

FAV_HD SC pixel_dc(const AltRenderParams& p, int px, int py) {  // renderer.cpp:12148-12152
    if (p.point_re) {  // point render: the caller's normalised offsets
        const size_t i = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
        return SC{p.point_re[i], p.point_im[i], p.point_e[i]};
    }
    return sc_from_fe(FE{p.x_off_m[px], p.x_off_e[px]}, FE{p.y_off_m[py], p.y_off_e[py]});
}

// The field value and escape count of one pixel (max_iter inside, -1 for a
// NaN pixel).
struct Result {
    float value;
    int count;
};

// The loop of render_alternate_reference_impl from a loop top: the
// perturbation `delta` against reference index `ref` at `iteration`, with
// Z + delta and its norm; dc, the parameter perturbation and its squared
// norm.  Counts the loop passes (blocks, exact and replayed steps) in `work`.
FAV_HD Result run_loop(const AltRenderParams& p, SC dc, SC pdelta, FE pn, SC delta, int ref, int iteration, SC total,
                       SN total_norm, int& work) {
    const bool julia = p.formula == 1;
    const int mode = p.escape_radius_mode;
    const double bias = p.output_bias;
    AltBudget budget{&p, 0U};
    const SN limit{0.75, 2};
    SC tortoise = sc_zero();
    int cycle_power = 1, cycle_length = 0, cycle_hits = 0;
    bool cycle_ready = false;
    while (iteration < p.max_iter) {
        if (p.work_limit > 0 && work >= p.work_limit) {
            return Result{__builtin_nanf(""), -1};
        }
        ++work;
        if (budget.expired()) {
            return Result{__builtin_nanf(""), -1};
        }
        const int next = ref + 1;
        if (next >= p.orbit_size) {  // past the reference's escape: exact recurrence
            const SC param{p.parameter_re, p.parameter_im, p.parameter_e};
            total = formula_step(p.formula, total, julia ? param : sc_add(param, dc));
            ++iteration;
            total_norm = sc_norm_squared(total);
            if (sc_outside_escape(total_norm, mode)) {
                return Result{core::smooth_escape(iteration, total_norm, bias), iteration};
            }
            continue;
        }
        const SN dn = sc_norm_squared(delta);
        const int remaining = p.max_iter - iteration;
        const int k = p.bla_enabled
            ? lookup(p, ref, FE{dn.m, dn.e}, pn, p.max_block < remaining ? p.max_block : remaining) : -1;
        if (k >= 0 && ref + p.steps[k].length < p.orbit_size) {
            const AltStepDev& step = p.steps[k];
            const int end = ref + step.length;
            const SC cand = apply(step, delta, pdelta);
            const FE cm = margin(p, end, cand);
            const SN cn = sn_from_margin(cm, mode);
            const SC ct = sc_add(load_z(p, end), cand);
            const bool bad = !sc_finite(cand) || !fe_finite(cm) || !sc_finite(ct) || fe_compare(cm, fe_zero()) > 0
                || sc_compare_norm(cn, limit) >= 0;
            if (!bad) {
                delta = cand;
                ref = end;
                iteration += step.length;
                total = ct;
                total_norm = cn;
                continue;
            }
            // replay the block one exact step at a time
            for (int replay = 0; replay < step.length && iteration < p.max_iter; ++replay) {
                if (p.work_limit > 0 && work >= p.work_limit) {
                    return Result{__builtin_nanf(""), -1};
                }
                ++work;
                delta = delta_step(p, ref, delta, pdelta);
                ++ref;
                ++iteration;
                if (!sc_finite(delta)) {
                    return Result{p.strict ? __builtin_nanf("") : core::encode_iteration(iteration, bias), p.strict ? -1 : iteration};
                }
                const FE m = margin(p, ref, delta);
                total_norm = sn_from_margin(m, mode);
                total = sc_add(load_z(p, ref), delta);
                if (fe_compare(m, fe_zero()) > 0 || sc_outside_escape(total_norm, mode)) {
                    return Result{core::smooth_escape(iteration, total_norm, bias), iteration};
                }
            }
            continue;
        }
        delta = delta_step(p, ref, delta, pdelta);
        ref = next;
        ++iteration;
        if (!sc_finite(delta)) {
            return Result{p.strict ? __builtin_nanf("") : core::encode_iteration(iteration, bias), p.strict ? -1 : iteration};
        }
        const FE m = margin(p, ref, delta);
        total_norm = sn_from_margin(m, mode);
        total = sc_add(load_z(p, ref), delta);
        if (fe_compare(m, fe_zero()) > 0 || sc_outside_escape(total_norm, mode)) {
            return Result{core::smooth_escape(iteration, total_norm, bias), iteration};
        }
        if (p.cycle_enabled && iteration >= 2048 && (iteration & 31) == 0 && sc_compare_norm(total_norm, limit) < 0) {
            if (!cycle_ready) {
                tortoise = total;
                cycle_power = 1;
                cycle_length = 0;
                cycle_hits = 0;
                cycle_ready = true;
            } else {
                const SN distance = sc_norm_squared(sc_sub(total, tortoise));
                const int scale = total_norm.e > 0 ? total_norm.e : 0;
                if (distance.m == 0.0 || distance.e <= scale - 78) {
                    if (++cycle_hits >= 3) {
                        return Result{core::encode_iteration(p.max_iter, bias), p.max_iter};
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
    return Result{core::encode_iteration(p.max_iter, bias), p.max_iter};
}


// One pixel of render_alternate_reference_impl; writes p.output[index] and
// its escape count, and counts the loop passes in `work`.
FAV_HD void pixel_body(const AltRenderParams& p, int px, int py, int& work, int& count) {
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    const bool julia = p.formula == 1;
    const int mode = p.escape_radius_mode;
    const double bias = p.output_bias;
    // the field value and the escape count (max_iter inside, -1 when NaN)
    auto finish = [&](float value, int n) {
        p.output[index] = value;
        count = n;
    };
    const SC dc = pixel_dc(p, px, py);
    const SC pdelta = julia ? sc_zero() : dc;
    const SN pnorm = sc_norm_squared(pdelta);
    const FE pn{pnorm.m, pnorm.e};
    SC delta = dc;
    int ref = julia ? 0 : 1, iteration = julia ? 0 : 1;
    if (p.series_available && sc_compare_norm(sc_norm_squared(dc), SN{p.series.radius2_m, p.series.radius2_e}) <= 0) {
        delta = series_value(p.series, dc);
        ref = p.series.iteration;
        iteration = p.series.iteration;
    }
    SC total = sc_add(load_z(p, ref), delta);
    SN total_norm = sn_from_margin(margin(p, ref, delta), mode);
    if (sc_outside_escape(total_norm, mode)) {
        finish(core::smooth_escape(iteration, total_norm, bias), iteration);
        return;
    }
    const Result r = run_loop(p, dc, pdelta, pn, delta, ref, iteration, total, total_norm, work);
    finish(r.value, r.count);
}

FAV_HD void pixel(const AltRenderParams& p, int px, int py) {
    int work = 0, count = 0;
    pixel_body(p, px, py, work, count);
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px);
    if (p.work) p.work[index] = work;
    if (p.counts) p.counts[index] = count;
}

}  // namespace alt
}  // namespace fav
