// fav_numeric.h -- extended-exponent numbers shared by the host table
// builders and the gfx950 field kernels.
//
// FloatExp (FE) is a double mantissa in [0.5, 1) with an int exponent;
// ScaledComplex (SC) is a complex pair sharing one exponent.  The operations
// below reproduce the rounding points of the reference's FloatExp /
// ScaledComplex (renderer.cpp:7532-7824, 8190-8474) exactly: every sum and
// product is rounded separately (the build uses -ffp-contract=off on both
// host and device), renormalisation takes the same fast paths, and the
// saturation and zero conventions match.  Escape counts of chaotic pixels
// depend on every one of those roundings.
//
// The same source compiles as host C++ (g++) and as HIP device code.
#pragma once

#include <climits>
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FAV_HD __host__ __device__ __forceinline__
#else
#define FAV_HD inline
#endif

namespace fav {

// ---- primitive scaling -------------------------------------------------
// frexp with the C library's conventions (0, +-inf and NaN return the input
// with exponent 0), split into its two results.  On the device these are
// v_frexp_exp_i32_f64 / v_frexp_mant_f64 / v_ldexp_f64, exact for every
// finite input including subnormals.
FAV_HD int frexp_exp_(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (v == 0.0 || !__builtin_isfinite(v)) ? 0 : __builtin_amdgcn_frexp_exp(v);
#else
    int e = 0;
    (void)std::frexp(v, &e);
    return e;
#endif
}

FAV_HD double frexp_(double v, int* e) {
#if defined(__HIP_DEVICE_COMPILE__)
    const bool ordinary = v != 0.0 && __builtin_isfinite(v);
    *e = ordinary ? __builtin_amdgcn_frexp_exp(v) : 0;
    return ordinary ? __builtin_amdgcn_frexp_mant(v) : v;
#else
    return std::frexp(v, e);
#endif
}

// frexp's exponent of a finite value (0 for zero; no infinity/NaN handling).
FAV_HD int frexp_exp_nz(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_frexp_exp(v);
#else
    int e = 0;
    (void)std::frexp(v, &e);
    return e;
#endif
}

FAV_HD double ldexp_(double v, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ldexp(v, e);
#else
    return std::ldexp(v, e);
#endif
}

FAV_HD bool isfinite_(double v) { return __builtin_isfinite(v); }
FAV_HD bool isnan_(double v) { return __builtin_isnan(v); }
FAV_HD double fabs_(double v) { return __builtin_fabs(v); }

// ---- saturating exponent arithmetic ---------------------------------------
// The reference forms exponent sums in long long and clamps them to int
// (saturating_exponent, renderer.cpp:7532-7540).  These give the same
// results with 32-bit operations.
FAV_HD int sat_exp(long long v) {
    return v > INT_MAX ? INT_MAX : (v < INT_MIN ? INT_MIN : static_cast<int>(v));
}

// clamp(a + b)
FAV_HD int sat_add(int a, int b) {
#if defined(__clang__)
    return __builtin_elementwise_add_sat(a, b);
#else
    int r;
    return __builtin_add_overflow(a, b, &r) ? (b > 0 ? INT_MAX : INT_MIN) : r;
#endif
}

// clamp(a + b + s) for s in {-1, 0, 1}: one saturating add of two exact
// operands, choosing the operand that absorbs s without overflow.
FAV_HD int sat_add3(int a, int b, int s) {
    const bool b_edge = (s > 0 && b == INT_MAX) || (s < 0 && b == INT_MIN);
    const bool a_edge = (s > 0 && a == INT_MAX) || (s < 0 && a == INT_MIN);
    if (!b_edge) return sat_add(a, b + s);
    if (!a_edge) return sat_add(a + s, b);
    return s > 0 ? INT_MAX : INT_MIN;
}

// ---- FloatExp ------------------------------------------------------------
struct FE {
    double m;
    int e;
};

FAV_HD FE fe(double m, int e) { return FE{m, e}; }
FAV_HD FE fe_zero() { return FE{0.0, 0}; }
FAV_HD bool fe_is_zero(FE v) { return v.m == 0.0; }
FAV_HD bool fe_finite(FE v) { return isfinite_(v.m); }

// FloatExp::from_parts, renderer.cpp:7551-7560.  Zero and NaN keep
// exponent 0; infinity keeps the given exponent (frexp leaves it alone).
FAV_HD FE fe_from_parts(double value, int exponent) {
    int shift = 0;
    const double n = frexp_(value, &shift);
    const FE ordinary{n, sat_add(exponent, shift)};
    return (value == 0.0 || isnan_(value)) ? FE{value, 0} : ordinary;
}

// escape_radius_squared_float_exp, renderer.cpp:7594-7598: 4 = 0.5*2^3,
// 1e8 = 0.7450580596923828125*2^27.
FAV_HD FE fe_escape_r2(int mode) {
    return mode == 1 ? FE{0.7450580596923828125, 27} : FE{0.5, 3};
}

// R^2 of an escape-radius mode as a plain double (4 or 1e8).
FAV_HD double escape_r2_double(int mode) { return mode == 1 ? 1.0e8 : 4.0; }

FAV_HD FE fe_neg(FE v) { return FE{-v.m, v.e}; }

// fe_add, renderer.cpp:8198-8235.  The reference renormalises the aligned
// sum by x0.5 when |m| >= 1, keeps it when |m| >= 0.5, and calls frexp
// otherwise; all three are ldexp(m, -shift) with the shift chosen below.
// The ordinary result is formed unconditionally and the special cases are
// selected over it in the reference's order of precedence, so the code has
// no divergent branches.
FAV_HD FE fe_add(FE a, FE b) {
    const bool swap = b.e > a.e;
    const FE larger = swap ? b : a;
    const FE smaller = swap ? a : b;
    const unsigned diff = static_cast<unsigned>(larger.e) - static_cast<unsigned>(smaller.e);
    const int align = static_cast<int>(diff < 64U ? diff : 64U);
    const double mantissa = larger.m + ldexp_(smaller.m, -align);
    const double magnitude = fabs_(mantissa);
    const int shift = magnitude >= 1.0 ? 1 : (magnitude >= 0.5 ? 0 : frexp_exp_(mantissa));
    FE r{ldexp_(mantissa, -shift), sat_add(larger.e, shift)};
    if (!isfinite_(mantissa)) r = FE{mantissa, larger.e};          // from_parts(+-inf, e)
    if (mantissa == 0.0 || isnan_(mantissa)) r = FE{mantissa, 0};  // from_parts(0 or NaN, e)
    if (diff > 60U) r = larger;
    if (b.m == 0.0) r = a;
    if (a.m == 0.0) r = b;
    return r;
}

FAV_HD FE fe_sub(FE a, FE b) { return fe_add(a, fe_neg(b)); }

// fe_mul, renderer.cpp:8241-8266: one x0.5 or x2 correction of the product.
FAV_HD FE fe_mul(FE a, FE b) {
    const double mantissa = a.m * b.m;
    const double magnitude = fabs_(mantissa);
    const int shift = magnitude >= 1.0 ? 1 : (magnitude < 0.5 ? -1 : 0);
    FE r{ldexp_(mantissa, -shift), sat_add3(a.e, b.e, shift)};
    if (!isfinite_(mantissa)) r = FE{mantissa, 0};  // from_parts(inf or NaN, 0)
    if (a.m == 0.0 || b.m == 0.0) r = FE{0.0, 0};
    return r;
}

// fe_mul(FloatExp, double), renderer.cpp:8268-8270.
FAV_HD FE fe_muld(FE a, double b) { return fe_mul(a, fe_from_parts(b, 0)); }

// fe_div, renderer.cpp:8272-8293.
FAV_HD FE fe_div(FE a, FE b) {
    if (a.m == 0.0) return FE{0.0, 0};
    const double mantissa = a.m / b.m;
    if (!isfinite_(mantissa)) return fe_from_parts(mantissa, 0);
    const double magnitude = fabs_(mantissa);
    const long long diff = static_cast<long long>(a.e) - b.e;
    if (magnitude >= 1.0) return FE{mantissa * 0.5, sat_exp(diff + 1)};
    if (magnitude < 0.5) return FE{mantissa * 2.0, sat_exp(diff - 1)};
    return FE{mantissa, sat_exp(diff)};
}

FAV_HD FE fe_sqr(FE v) { return fe_mul(v, v); }

// fe_sqrt, renderer.cpp:8299-8316.
FAV_HD FE fe_sqrt(FE v) {
    if (v.m == 0.0) return v;
    long long exponent = v.e;
    double mantissa = v.m;
    if (mantissa < 0.0 || !isfinite_(mantissa)) return fe_from_parts(std::sqrt(mantissa), 0);
    if (exponent % 2 != 0) {
        mantissa *= 2.0;
        --exponent;
        return FE{std::sqrt(mantissa) * 0.5, sat_exp(exponent / 2 + 1)};
    }
    return FE{std::sqrt(mantissa), sat_exp(exponent / 2)};
}

// fe_compare, renderer.cpp:8318-8339.
FAV_HD int fe_compare(FE a, FE b) {
    if (!isfinite_(a.m)) return isfinite_(b.m) ? 1 : 0;
    if (!isfinite_(b.m)) return -1;
    if (a.m == 0.0) return b.m == 0.0 ? 0 : -1;
    if (b.m == 0.0) return a.m > 0.0 ? 1 : -1;
    if (a.m < 0.0 && b.m >= 0.0) return -1;
    if (a.m >= 0.0 && b.m < 0.0) return 1;
    const bool negative = a.m < 0.0;
    if (a.e != b.e) {
        const int result = a.e < b.e ? -1 : 1;
        return negative ? -result : result;
    }
    return (a.m > b.m) - (a.m < b.m);
}

// fe_compare(a, b) < 0 for finite, non-negative b (validity radii and
// glitch thresholds).
FAV_HD bool fe_less_nonneg(FE a, FE b) {
    if (!isfinite_(a.m)) return false;
    if (a.m == 0.0) return b.m != 0.0;
    if (b.m == 0.0) return a.m < 0.0;
    if (a.m < 0.0) return true;
    return a.e != b.e ? a.e < b.e : a.m < b.m;
}

FAV_HD FE fe_min(FE a, FE b) { return fe_compare(a, b) <= 0 ? a : b; }  // renderer.cpp:11336

// ---- FloatExpComplex (renderer.cpp:8351-8399) -----------------------------
struct FEC {
    FE re, im;
};

FAV_HD FEC fec(FE re, FE im) { return FEC{re, im}; }
FAV_HD FEC fec_add(FEC a, FEC b) { return FEC{fe_add(a.re, b.re), fe_add(a.im, b.im)}; }
FAV_HD FEC fec_neg(FEC v) { return FEC{fe_neg(v.re), fe_neg(v.im)}; }
FAV_HD FEC fec_sub(FEC a, FEC b) { return fec_add(a, fec_neg(b)); }
FAV_HD FEC fec_mul(FEC a, FEC b) {
    return FEC{fe_sub(fe_mul(a.re, b.re), fe_mul(a.im, b.im)),
               fe_add(fe_mul(a.re, b.im), fe_mul(a.im, b.re))};
}
FAV_HD FEC fec_mulf(FEC a, FE b) { return FEC{fe_mul(a.re, b), fe_mul(a.im, b)}; }
FAV_HD FE fec_norm_squared(FEC v) { return fe_add(fe_sqr(v.re), fe_sqr(v.im)); }
FAV_HD bool fec_finite(FEC v) { return fe_finite(v.re) && fe_finite(v.im); }

// fec_escape_margin_with_delta, renderer.cpp:8383-8399:
// |Z+d|^2 - R^2 = (|Z|^2 - R^2) + 2 Re(Z conj d) + |d|^2.
FAV_HD FE fec_margin_from_reference(FE reference_margin, FEC reference, FEC delta) {
    const FE cross = fe_muld(fe_add(fe_mul(reference.re, delta.re),
                                    fe_mul(reference.im, delta.im)), 2.0);
    return fe_add(fe_add(reference_margin, cross), fec_norm_squared(delta));
}

FAV_HD FE fec_escape_margin_with_delta(FEC reference, FEC delta, int mode) {
    const FE reference_margin = fe_sub(fec_norm_squared(reference), fe_escape_r2(mode));
    return fec_margin_from_reference(reference_margin, reference, delta);
}

// ---- ScaledComplex (renderer.cpp:7605-7736) -------------------------------
struct SC {
    double re, im;
    int e;
};

struct SN {  // ScaledNorm
    double m;
    int e;
};

FAV_HD SC sc(double re, double im, int e) { return SC{re, im, e}; }
FAV_HD SC sc_zero() { return SC{0.0, 0.0, 0}; }
FAV_HD bool sc_is_zero(SC v) { return v.re == 0.0 && v.im == 0.0; }

// ScaledComplex::normalize, renderer.cpp:7610-7655.  Each of the
// reference's cases (frexp above 2, x0.5 in [1,2), none in [0.5,1), x2 in
// [0.25,0.5), frexp below) scales both parts by 2^-shift with shift equal to
// frexp's exponent of the larger magnitude, so that single shift reproduces
// every case bit for bit; non-finite magnitudes keep shift 0.
FAV_HD void sc_normalize(SC& v) {
    const double ar = fabs_(v.re), ai = fabs_(v.im);
    const double magnitude = (ar < ai) ? ai : ar;  // std::max(ar, ai)
    const int shift = frexp_exp_(magnitude);
    const bool zero = magnitude == 0.0;
    v.re = zero ? 0.0 : ldexp_(v.re, -shift);
    v.im = zero ? 0.0 : ldexp_(v.im, -shift);
    v.e = zero ? 0 : sat_add(v.e, shift);
}

// ScaledComplex::from_float_exp, renderer.cpp:7657-7674.
FAV_HD SC sc_from_fe(FE re, FE im) {
    if (re.m == 0.0 && im.m == 0.0) return SC{0.0, 0.0, 0};
    const int common = re.e > im.e ? re.e : im.e;
    const long long rd = static_cast<long long>(re.e) - common;
    const long long id = static_cast<long long>(im.e) - common;
    SC r{rd < -1074 ? 0.0 : ldexp_(re.m, static_cast<int>(rd)),
         id < -1074 ? 0.0 : ldexp_(im.m, static_cast<int>(id)), common};
    sc_normalize(r);
    return r;
}

// sc_add, renderer.cpp:7676-7697 (special cases selected, no branches).
FAV_HD SC sc_add(SC a, SC b) {
    const bool swap = b.e > a.e;
    const SC larger = swap ? b : a;
    const SC smaller = swap ? a : b;
    const unsigned diff = static_cast<unsigned>(larger.e) - static_cast<unsigned>(smaller.e);
    const int d = static_cast<int>(diff < 64U ? diff : 64U);
    SC r{larger.re + ldexp_(smaller.re, -d), larger.im + ldexp_(smaller.im, -d), larger.e};
    sc_normalize(r);
    if (diff > 60U) r = larger;
    if (b.re == 0.0 && b.im == 0.0) r = a;
    if (a.re == 0.0 && a.im == 0.0) r = b;
    return r;
}

FAV_HD SC sc_neg(SC v) { return SC{-v.re, -v.im, v.e}; }
FAV_HD SC sc_conj(SC v) { return SC{v.re, -v.im, v.e}; }
FAV_HD SC sc_sub(SC a, SC b) { return sc_add(a, sc_neg(b)); }

// sc_mul, renderer.cpp:7711-7722 (no fused multiply-add).
FAV_HD SC sc_mul(SC a, SC b) {
    SC r{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re, sat_add(a.e, b.e)};
    sc_normalize(r);
    if ((a.re == 0.0 && a.im == 0.0) || (b.re == 0.0 && b.im == 0.0)) r = SC{0.0, 0.0, 0};
    return r;
}

// sc_double, renderer.cpp:7724-7733: exact, no renormalisation.
FAV_HD SC sc_double(SC v) {
    if (v.re == 0.0 && v.im == 0.0) return SC{0.0, 0.0, 0};
    return SC{v.re, v.im, sat_add(v.e, 1)};
}

FAV_HD bool sc_finite(SC v) { return isfinite_(v.re) && isfinite_(v.im); }

// ---- operations on normalised nonzero operands ---------------------------
// A ScaledComplex is "normal" when its larger part has magnitude in
// [0.5, 1) and |e| <= 2^28: what normalize() leaves for every finite
// nonzero value in the range the field path reaches.  For normal operands
// the special cases of sc_mul / sc_add / sc_norm_squared cannot fire:
//  * a product of two normal values has magnitude in [0.25, 2), so its
//    larger part is in [0.17, 2) (each part is rounded once, so cancellation
//    in one part leaves the other near |a||b|) -- never zero or infinite;
//  * a sum of two normal values is below 2 in magnitude but may cancel to
//    exactly zero, which normalize() maps to exponent 0;
//  * |v|^2 of a normal value lies in [0.25, 2);
//  * exponent sums stay far from the int limits, so no saturation.
// The *_nn forms below compute the same roundings in the same order with
// those cases removed, so for normal operands they equal sc_mul, sc_add and
// sc_norm_squared bit for bit.
FAV_HD bool sc_normal(SC v) {
    const double ar = fabs_(v.re), ai = fabs_(v.im);
    const double magnitude = (ar < ai) ? ai : ar;
    return magnitude >= 0.5 && magnitude < 1.0 && v.e <= (1 << 28) && v.e >= -(1 << 28);
}

// max(|x|, |y|) of finite values: the larger magnitude, as the reference's
// (|x| < |y|) ? |y| : |x| selects it (one instruction on the device).
FAV_HD double max_abs(double x, double y) { return __builtin_fmax(fabs_(x), fabs_(y)); }

FAV_HD SC sc_mul_nn(SC a, SC b) {
    const double re = a.re * b.re - a.im * b.im;
    const double im = a.re * b.im + a.im * b.re;
    const int shift = frexp_exp_nz(max_abs(re, im));
    return SC{ldexp_(re, -shift), ldexp_(im, -shift), a.e + b.e + shift};
}

// Both operands are aligned to the larger exponent (the larger one by a
// shift of 0), which gives the same sums as adding the shifted smaller
// operand to the larger (IEEE addition is commutative) without swapping.
FAV_HD SC sc_add_nn(SC a, SC b) {
    const int top = a.e > b.e ? a.e : b.e;
    const int da = top - a.e < 64 ? top - a.e : 64;
    const int db = top - b.e < 64 ? top - b.e : 64;
    const double re = ldexp_(a.re, -da) + ldexp_(b.re, -db);
    const double im = ldexp_(a.im, -da) + ldexp_(b.im, -db);
    const double magnitude = max_abs(re, im);
    const int shift = frexp_exp_nz(magnitude);  // 0 for an exact cancellation
    SC r{ldexp_(re, -shift), ldexp_(im, -shift), magnitude == 0.0 ? 0 : top + shift};
    if (da + db > 60) r = b.e > a.e ? b : a;  // the larger operand unchanged
    return r;
}

// |v|^2 of a normal value: the sum lies in [0.25, 2), so the frexp exponent
// k is -1, 0 or 1 and scaling by 2^-k is the reference's *2, *1 or *0.5.
FAV_HD SN sc_norm_nn(SC v) {
    const double norm = v.re * v.re + v.im * v.im;
    const int k = frexp_exp_nz(norm);
    return SN{ldexp_(norm, -k), 2 * v.e + k};
}

// sc_compare_norm(a, b) >= 0 for a finite-or-not `a` and a normal, nonzero,
// finite `b` (renderer.cpp:7776-7784 restricted to that case).
FAV_HD bool sn_at_least(SN a, SN b) {
    if (!isfinite_(a.m)) return true;
    if (a.m == 0.0) return false;
    return a.e != b.e ? a.e > b.e : a.m >= b.m;
}

// sc_norm_squared, renderer.cpp:7743-7774.
FAV_HD SN sc_norm_squared(SC v) {
    const SN inf{__builtin_inf(), INT_MAX};
    if (!sc_finite(v)) return inf;
    const double norm = v.re * v.re + v.im * v.im;
    if (norm == 0.0) return SN{0.0, 0};
    if (!isfinite_(norm)) return inf;
    const long long se = 2LL * v.e;
    if (norm >= 1.0) {
        const long long e = se + 1;
        if (e > INT_MAX) return inf;
        return SN{norm * 0.5, static_cast<int>(e)};
    }
    if (norm < 0.5) {
        const long long e = se - 1;
        if (e < INT_MIN) return SN{0.0, 0};
        return SN{norm * 2.0, static_cast<int>(e)};
    }
    if (se > INT_MAX) return inf;
    if (se < INT_MIN) return SN{0.0, 0};
    return SN{norm, static_cast<int>(se)};
}

// sc_compare_norm, renderer.cpp:7776-7784.
FAV_HD int sc_compare_norm(SN a, SN b) {
    if (!isfinite_(a.m)) return isfinite_(b.m) ? 1 : 0;
    if (!isfinite_(b.m)) return -1;
    if (a.m == 0.0 && b.m == 0.0) return 0;
    if (a.m == 0.0) return -1;
    if (b.m == 0.0) return a.m > 0.0 ? 1 : -1;
    if (a.e != b.e) return a.e < b.e ? -1 : 1;
    return (a.m > b.m) - (a.m < b.m);
}

FAV_HD bool sc_outside_escape(SN n, int mode) {  // renderer.cpp:7786-7793
    const FE t = fe_escape_r2(mode);
    return sc_compare_norm(n, SN{t.m, t.e}) > 0;
}

FAV_HD double sc_re_double(SC v) { return ldexp_(v.re, v.e); }
FAV_HD double sc_im_double(SC v) { return ldexp_(v.im, v.e); }

// sc_component_as_float_exp, renderer.cpp:8401-8407.
FAV_HD FEC sc_parts(SC v) { return FEC{fe_from_parts(v.re, v.e), fe_from_parts(v.im, v.e)}; }

// sc_norm_squared_with_delta, renderer.cpp:8451-8461.
FAV_HD SN sn_from_margin(FE margin, int mode) {
    const FE n = fe_add(fe_escape_r2(mode), margin);
    return SN{n.m, n.e};
}

// scaled_norm_to_double / scaled_norm_from_double, renderer.cpp:7803-7824.
FAV_HD double sn_to_double(SN n) {
    if (n.m == 0.0) return 0.0;
    if (!isfinite_(n.m)) return 1.7976931348623157e308;
    const double value = ldexp_(n.m, n.e);
    return isfinite_(value) ? (0.0 < value ? value : 0.0) : 1.7976931348623157e308;
}

FAV_HD SN sn_from_double(double value) {
    if (!(value > 0.0) || !isfinite_(value))
        return SN{value > 0.0 ? 1.7976931348623157e308 : 0.0, 0};
    int e = 0;
    const double m = frexp_(value, &e);
    return SN{m, e};
}

}  // namespace fav
