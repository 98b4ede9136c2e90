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
// frexp/ldexp with the C library's conventions (0, +-inf and NaN return the
// input with exponent 0).  On the device they map onto v_frexp_mant_f64,
// v_frexp_exp_i32_f64 and v_ldexp_f64, which are exact for every finite
// input including subnormals.
FAV_HD double frexp_(double v, int* e) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (v == 0.0 || !__builtin_isfinite(v)) { *e = 0; return v; }
    *e = __builtin_amdgcn_frexp_exp(v);
    return __builtin_amdgcn_frexp_mant(v);
#else
    return std::frexp(v, e);
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

FAV_HD int sat_exp(long long v) {  // renderer.cpp:7532-7540
    return v > INT_MAX ? INT_MAX : (v < INT_MIN ? INT_MIN : static_cast<int>(v));
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

// FloatExp::from_parts, renderer.cpp:7551-7560.
FAV_HD FE fe_from_parts(double value, int exponent) {
    if (value == 0.0 || isnan_(value)) return FE{value, 0};
    int shift = 0;
    const double n = frexp_(value, &shift);
    return FE{n, sat_exp(static_cast<long long>(exponent) + shift)};
}

// escape_radius_squared_float_exp, renderer.cpp:7594-7598: 4 = 0.5*2^3,
// 1e8 = 0.7450580596923828*2^27.
FAV_HD FE fe_escape_r2(int mode) {
    return mode == 1 ? FE{0.7450580596923828125, 27} : FE{0.5, 3};
}

FAV_HD FE fe_neg(FE v) { return FE{-v.m, v.e}; }

// fe_add, renderer.cpp:8198-8235.
FAV_HD FE fe_add(FE a, FE b) {
    if (a.m == 0.0) return b;
    if (b.m == 0.0) return a;
    FE larger = a, smaller = b;
    if (b.e > a.e) { larger = b; smaller = a; }
    const long long diff = static_cast<long long>(larger.e) - smaller.e;
    if (diff > 60) return larger;
    const double mantissa = larger.m + ldexp_(smaller.m, -static_cast<int>(diff));
    if (mantissa == 0.0 || !isfinite_(mantissa)) return fe_from_parts(mantissa, larger.e);
    const double magnitude = fabs_(mantissa);
    if (magnitude >= 1.0) return FE{mantissa * 0.5, sat_exp(static_cast<long long>(larger.e) + 1)};
    if (magnitude >= 0.5) return FE{mantissa, larger.e};
    int shift = 0;
    const double n = frexp_(mantissa, &shift);
    return FE{n, sat_exp(static_cast<long long>(larger.e) + shift)};
}

FAV_HD FE fe_sub(FE a, FE b) { return fe_add(a, fe_neg(b)); }

// fe_mul, renderer.cpp:8241-8266.
FAV_HD FE fe_mul(FE a, FE b) {
    if (a.m == 0.0 || b.m == 0.0) return FE{0.0, 0};
    const double mantissa = a.m * b.m;
    if (!isfinite_(mantissa)) return fe_from_parts(mantissa, 0);
    const double magnitude = fabs_(mantissa);
    const long long sum = static_cast<long long>(a.e) + b.e;
    if (magnitude >= 1.0) return FE{mantissa * 0.5, sat_exp(sum + 1)};
    if (magnitude < 0.5) return FE{mantissa * 2.0, sat_exp(sum - 1)};
    return FE{mantissa, sat_exp(sum)};
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

// ScaledComplex::normalize, renderer.cpp:7610-7655.
FAV_HD void sc_normalize(SC& v) {
    const double ar = fabs_(v.re), ai = fabs_(v.im);
    const double magnitude = (ar < ai) ? ai : ar;  // std::max(ar, ai)
    if (magnitude == 0.0) { v.re = 0.0; v.im = 0.0; v.e = 0; return; }
    if (magnitude >= 2.0) {
        int shift = 0;
        (void)frexp_(magnitude, &shift);
        v.re = ldexp_(v.re, -shift);
        v.im = ldexp_(v.im, -shift);
        v.e = sat_exp(static_cast<long long>(v.e) + shift);
        return;
    }
    if (magnitude >= 1.0) {
        v.re *= 0.5; v.im *= 0.5;
        v.e = sat_exp(static_cast<long long>(v.e) + 1);
        return;
    }
    if (magnitude >= 0.5) return;
    if (magnitude >= 0.25) {
        v.re *= 2.0; v.im *= 2.0;
        v.e = sat_exp(static_cast<long long>(v.e) - 1);
        return;
    }
    int shift = 0;
    (void)frexp_(magnitude, &shift);
    v.re = ldexp_(v.re, -shift);
    v.im = ldexp_(v.im, -shift);
    v.e = sat_exp(static_cast<long long>(v.e) + shift);
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

// sc_add, renderer.cpp:7676-7697.
FAV_HD SC sc_add(SC a, SC b) {
    if (a.re == 0.0 && a.im == 0.0) return b;
    if (b.re == 0.0 && b.im == 0.0) return a;
    SC larger = a, smaller = b;
    if (b.e > a.e) { larger = b; smaller = a; }
    const long long diff = static_cast<long long>(larger.e) - smaller.e;
    if (diff > 60) return larger;
    const int d = static_cast<int>(diff);
    SC r{larger.re + ldexp_(smaller.re, -d), larger.im + ldexp_(smaller.im, -d), larger.e};
    sc_normalize(r);
    return r;
}

FAV_HD SC sc_neg(SC v) { return SC{-v.re, -v.im, v.e}; }
FAV_HD SC sc_conj(SC v) { return SC{v.re, -v.im, v.e}; }
FAV_HD SC sc_sub(SC a, SC b) { return sc_add(a, sc_neg(b)); }

// sc_mul, renderer.cpp:7711-7722 (no fused multiply-add).
FAV_HD SC sc_mul(SC a, SC b) {
    if ((a.re == 0.0 && a.im == 0.0) || (b.re == 0.0 && b.im == 0.0)) return SC{0.0, 0.0, 0};
    SC r{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re,
         sat_exp(static_cast<long long>(a.e) + b.e)};
    sc_normalize(r);
    return r;
}

// sc_double, renderer.cpp:7724-7733: exact, no renormalisation.
FAV_HD SC sc_double(SC v) {
    if (v.re == 0.0 && v.im == 0.0) return SC{0.0, 0.0, 0};
    return SC{v.re, v.im, sat_exp(static_cast<long long>(v.e) + 1)};
}

FAV_HD bool sc_finite(SC v) { return isfinite_(v.re) && isfinite_(v.im); }

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
