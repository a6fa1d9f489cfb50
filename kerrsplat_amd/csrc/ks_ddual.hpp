// Scalar types of the device transfer code, which is written once as templates on its scalar:
//   * double     - the plain FP64 path;
//   * f32        - the FP32 path: a float whose arithmetic with double constants stays in float
//                  (the fits and the step are written with double literals; with f32 every
//                  operation rounds to float and runs on the FP32 units and hardware transcendentals);
//   * DN<N, S>   - forward-mode duals (a value and N partials of scalar S) for the per-sample local
//                  derivatives of the adjoint, as the reference obtains them with ForwardDiff duals
//                  (src/Splats/dual_gradient.jl:1-19, src/Transfer/transport.jl:198-207).
#pragma once
#include <type_traits>
#include "ks_dmath.hpp"

namespace ksd {

// the double overloads stay visible beside the ones declared in this namespace
using ::sqrt; using ::exp; using ::expm1; using ::log; using ::log1p; using ::sin; using ::cos; using ::sincos;
using ::sinh; using ::cosh; using ::tanh; using ::cbrt; using ::acos; using ::fabs; using ::atan2; using ::pow; using ::frexp;
using ::atan; using ::floor;

// ---- f32 ------------------------------------------------------------------------------------------------
struct f32 {
    float v;
    KS_DEV f32() {}
    KS_DEV f32(float x) : v(x) {}
    KS_DEV f32(double x) : v((float)x) {}
    KS_DEV f32(int x) : v((float)x) {}
};
KS_DEV double val(double x) { return x; }
KS_DEV double val(f32 x) { return (double)x.v; }

#define KSF_BIN(op)                                                                                \
    KS_DEV f32 operator op(f32 a, f32 b) { return f32(a.v op b.v); }                               \
    KS_DEV f32 operator op(f32 a, double b) { return f32(a.v op (float)b); }                       \
    KS_DEV f32 operator op(double a, f32 b) { return f32((float)a op b.v); }
KSF_BIN(+)
KSF_BIN(-)
KSF_BIN(*)
KSF_BIN(/)
#undef KSF_BIN
#define KSF_CMP(op)                                                                                \
    KS_DEV bool operator op(f32 a, f32 b) { return a.v op b.v; }                                   \
    KS_DEV bool operator op(f32 a, double b) { return a.v op (float)b; }                           \
    KS_DEV bool operator op(double a, f32 b) { return (float)a op b.v; }
KSF_CMP(<)
KSF_CMP(>)
KSF_CMP(<=)
KSF_CMP(>=)
KSF_CMP(==)
KSF_CMP(!=)
#undef KSF_CMP
KS_DEV f32 operator-(f32 a) { return f32(-a.v); }
KS_DEV f32& operator+=(f32& a, f32 b) { a.v += b.v; return a; }
KS_DEV f32& operator-=(f32& a, f32 b) { a.v -= b.v; return a; }
KS_DEV f32& operator*=(f32& a, f32 b) { a.v *= b.v; return a; }
KS_DEV f32 sqrt(f32 x) { return f32(sqrtf(x.v)); }
// pow of the FP32 path on the hardware base-2 transcendentals (v_log_f32, v_exp_f32, about one ulp each):
// x^y = 2^(y log2 x) for x > 0, with a relative error that grows with the exponent's magnitude, |y log2 x| 2^-24
// (4e-6 at |y log2 x| = 100), far inside the FP32 path's error bars; a base that is zero, negative or not finite
// takes the library's powf. exp, log and cbrt stay the library's: exp through 2^(x log2 e) loses |x| 2^-24 of
// relative accuracy, which optically thick, Faraday-rotating parcels amplify past the FP32 path's bar.
// (host instantiations, the work model and host checks, take the C library's base-2 functions)
KS_DEV float hw_exp2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}
KS_DEV float hw_log2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}
KS_DEV float hw_pow(float x, float y) {
    if (x > 0.0f && x < INFINITY) return hw_exp2(y * hw_log2(x));
    return powf(x, y);
}
KS_DEV f32 exp(f32 x) { return f32(expf(x.v)); }
KS_DEV f32 expm1(f32 x) { return f32(expm1f(x.v)); }
KS_DEV f32 log(f32 x) { return f32(logf(x.v)); }
KS_DEV f32 log1p(f32 x) { return f32(log1pf(x.v)); }
KS_DEV f32 sin(f32 x) { return f32(sinf(x.v)); }
KS_DEV f32 cos(f32 x) { return f32(cosf(x.v)); }
KS_DEV void sincos(f32 x, f32* s, f32* c) { float sv, cv; sincosf(x.v, &sv, &cv); *s = f32(sv); *c = f32(cv); }
KS_DEV f32 sinh(f32 x) { return f32(sinhf(x.v)); }
KS_DEV f32 cosh(f32 x) { return f32(coshf(x.v)); }
KS_DEV f32 tanh(f32 x) { return f32(tanhf(x.v)); }
KS_DEV f32 cbrt(f32 x) { return f32(cbrtf(x.v)); }
KS_DEV f32 acos(f32 x) { return f32(acosf(x.v)); }
KS_DEV f32 fabs(f32 x) { return f32(fabsf(x.v)); }
KS_DEV f32 atan2(f32 y, f32 x) { return f32(atan2f(y.v, x.v)); }
KS_DEV f32 atan(f32 x) { return f32(atanf(x.v)); }
KS_DEV f32 floor(f32 x) { return f32(floorf(x.v)); }
KS_DEV f32 pow(f32 x, double y) { return f32(hw_pow(x.v, (float)y)); }
KS_DEV f32 pow(f32 x, f32 y) { return f32(hw_pow(x.v, y.v)); }
KS_DEV f32 pow(double x, f32 y) { return f32(hw_pow((float)x, y.v)); }
KS_DEV f32 dmax(f32 a, f32 b) { return a.v > b.v ? a : b; }
KS_DEV f32 dmin(f32 a, f32 b) { return a.v < b.v ? a : b; }
KS_DEV f32 dmax(f32 a, double b) { return a.v > (float)b ? a : f32(b); }
KS_DEV f32 dmin(f32 a, double b) { return a.v < (float)b ? a : f32(b); }
KS_DEV bool isfinite_(double x) { return isfinite(x); }
KS_DEV bool isfinite_(f32 x) { return isfinite(x.v); }

// ---- forward-mode duals over a base scalar S ---------------------------------------------------------------
template <int N, class S = double> struct DN {
    S v;
    S d[N];
    KS_DEV DN() {}
    KS_DEV DN(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = S(0.0);
    }
    template <class X = S, class = typename std::enable_if<!std::is_same<X, double>::value>::type>
    KS_DEV DN(S x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = S(0.0);
    }
    KS_DEV static DN seed(S x, int k) { DN r; r.v = x; for (int i = 0; i < N; ++i) r.d[i] = S(0.0); r.d[k] = S(1.0); return r; }
};
template <int N, class S> KS_DEV double val(const DN<N, S>& x) { return val(x.v); }

#define KSDD_T template <int N, class S>
#define KSDD_LOOP _Pragma("unroll") for (int i = 0; i < N; ++i)
KSDD_T KS_DEV DN<N, S> operator+(const DN<N, S>& a, const DN<N, S>& b) { DN<N, S> r; r.v = a.v + b.v; KSDD_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
KSDD_T KS_DEV DN<N, S> operator-(const DN<N, S>& a, const DN<N, S>& b) { DN<N, S> r; r.v = a.v - b.v; KSDD_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
KSDD_T KS_DEV DN<N, S> operator*(const DN<N, S>& a, const DN<N, S>& b) { DN<N, S> r; r.v = a.v * b.v; KSDD_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
// one reciprocal per dual division (the value then carries at most one more rounding)
KSDD_T KS_DEV DN<N, S> operator/(const DN<N, S>& a, const DN<N, S>& b) {
    DN<N, S> r; const S ib = 1.0 / b.v; r.v = a.v * ib;
    KSDD_LOOP r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
KSDD_T KS_DEV DN<N, S> operator+(const DN<N, S>& a, double b) { DN<N, S> r = a; r.v = r.v + b; return r; }
KSDD_T KS_DEV DN<N, S> operator+(double a, const DN<N, S>& b) { DN<N, S> r = b; r.v = r.v + a; return r; }
KSDD_T KS_DEV DN<N, S> operator-(const DN<N, S>& a, double b) { DN<N, S> r = a; r.v = r.v - b; return r; }
KSDD_T KS_DEV DN<N, S> operator-(double a, const DN<N, S>& b) { DN<N, S> r; r.v = a - b.v; KSDD_LOOP r.d[i] = -b.d[i]; return r; }
KSDD_T KS_DEV DN<N, S> operator*(const DN<N, S>& a, double b) { DN<N, S> r; r.v = a.v * b; KSDD_LOOP r.d[i] = a.d[i] * b; return r; }
KSDD_T KS_DEV DN<N, S> operator*(double a, const DN<N, S>& b) { return b * a; }
KSDD_T KS_DEV DN<N, S> operator/(const DN<N, S>& a, double b) { return a * (1.0 / b); }
KSDD_T KS_DEV DN<N, S> operator/(double a, const DN<N, S>& b) { DN<N, S> r; const S ib = 1.0 / b.v; r.v = a * ib; const S f = -r.v * ib; KSDD_LOOP r.d[i] = f * b.d[i]; return r; }
KSDD_T KS_DEV DN<N, S> operator-(const DN<N, S>& a) { DN<N, S> r; r.v = -a.v; KSDD_LOOP r.d[i] = -a.d[i]; return r; }
KSDD_T KS_DEV DN<N, S>& operator+=(DN<N, S>& a, const DN<N, S>& b) { a = a + b; return a; }
KSDD_T KS_DEV DN<N, S>& operator*=(DN<N, S>& a, const DN<N, S>& b) { a = a * b; return a; }
KSDD_T KS_DEV DN<N, S>& operator*=(DN<N, S>& a, double b) { a = a * b; return a; }
// duals times their base scalar when that is not double (double has its overloads above)
#define KSDD_TS template <int N, class S, class = typename std::enable_if<!std::is_same<S, double>::value>::type>
KSDD_TS KS_DEV DN<N, S> operator*(const DN<N, S>& a, S b) { DN<N, S> r; r.v = a.v * b; KSDD_LOOP r.d[i] = a.d[i] * b; return r; }
KSDD_TS KS_DEV DN<N, S> operator*(S a, const DN<N, S>& b) { DN<N, S> r; r.v = a * b.v; KSDD_LOOP r.d[i] = a * b.d[i]; return r; }
KSDD_TS KS_DEV DN<N, S> operator+(const DN<N, S>& a, S b) { DN<N, S> r = a; r.v = r.v + b; return r; }
KSDD_TS KS_DEV DN<N, S> operator+(S a, const DN<N, S>& b) { DN<N, S> r = b; r.v = r.v + a; return r; }
KSDD_TS KS_DEV DN<N, S> operator-(const DN<N, S>& a, S b) { DN<N, S> r = a; r.v = r.v - b; return r; }
KSDD_TS KS_DEV DN<N, S> operator-(S a, const DN<N, S>& b) { DN<N, S> r; r.v = a - b.v; KSDD_LOOP r.d[i] = -b.d[i]; return r; }
KSDD_TS KS_DEV DN<N, S> operator/(const DN<N, S>& a, S b) { const S ib = 1.0 / b; return a * ib; }
KSDD_TS KS_DEV DN<N, S> operator/(S a, const DN<N, S>& b) { DN<N, S> r; const S ib = 1.0 / b.v; r.v = a * ib; const S f = -r.v * ib; KSDD_LOOP r.d[i] = f * b.d[i]; return r; }
#undef KSDD_TS
#define KSDD_CMP(op)                                                                                     \
    KSDD_T KS_DEV bool operator op(const DN<N, S>& a, const DN<N, S>& b) { return a.v op b.v; }         \
    KSDD_T KS_DEV bool operator op(const DN<N, S>& a, double b) { return a.v op b; }                    \
    KSDD_T KS_DEV bool operator op(double a, const DN<N, S>& b) { return a op b.v; }
KSDD_CMP(<)
KSDD_CMP(>)
KSDD_CMP(<=)
KSDD_CMP(>=)
KSDD_CMP(==)
KSDD_CMP(!=)
#undef KSDD_CMP

template <class S> struct ident_t { using type = S; };
KSDD_T KS_DEV DN<N, S> chain(typename ident_t<S>::type f, typename ident_t<S>::type df, const DN<N, S>& x) {
    DN<N, S> r; r.v = f; KSDD_LOOP r.d[i] = df * x.d[i]; return r;
}
KSDD_T KS_DEV DN<N, S> sqrt(const DN<N, S>& x) { const S s = sqrt(x.v); return chain(s, s > 0.0 ? S(0.5 / s) : S(0.0), x); }
KSDD_T KS_DEV DN<N, S> exp(const DN<N, S>& x) { const S e = exp(x.v); return chain(e, e, x); }
KSDD_T KS_DEV DN<N, S> expm1(const DN<N, S>& x) { return chain(expm1(x.v), exp(x.v), x); }
KSDD_T KS_DEV DN<N, S> log(const DN<N, S>& x) { return chain(log(x.v), S(1.0 / x.v), x); }
KSDD_T KS_DEV DN<N, S> log1p(const DN<N, S>& x) { return chain(log1p(x.v), S(1.0 / (1.0 + x.v)), x); }
KSDD_T KS_DEV DN<N, S> sin(const DN<N, S>& x) { S s, c; sincos(x.v, &s, &c); return chain(s, c, x); }
KSDD_T KS_DEV DN<N, S> cos(const DN<N, S>& x) { S s, c; sincos(x.v, &s, &c); return chain(c, S(-s), x); }
KSDD_T KS_DEV void sincos(const DN<N, S>& x, DN<N, S>* s, DN<N, S>* c) {
    S sv, cv; sincos(x.v, &sv, &cv);
    *s = chain(sv, cv, x); *c = chain(cv, S(-sv), x);
}
KSDD_T KS_DEV DN<N, S> sinh(const DN<N, S>& x) { return chain(sinh(x.v), cosh(x.v), x); }
KSDD_T KS_DEV DN<N, S> cosh(const DN<N, S>& x) { return chain(cosh(x.v), sinh(x.v), x); }
KSDD_T KS_DEV DN<N, S> tanh(const DN<N, S>& x) { const S t = tanh(x.v); return chain(t, S(1.0 - t * t), x); }
KSDD_T KS_DEV DN<N, S> cbrt(const DN<N, S>& x) { const S c = cbrt(x.v); return chain(c, S(c / (3.0 * x.v)), x); }
// acos with a zero derivative at the ends (the reference's safe_acos, step.jl:285-291)
KSDD_T KS_DEV DN<N, S> acos(const DN<N, S>& x) {
    const S s = 1.0 - x.v * x.v;
    return chain(acos(x.v), s > 0.0 ? S(-1.0 / sqrt(s)) : S(0.0), x);
}
KSDD_T KS_DEV DN<N, S> fabs(const DN<N, S>& x) { return x.v < 0.0 ? -x : x; }
// (the base-scalar calls are qualified)
KSDD_T KS_DEV DN<N, S> atan(const DN<N, S>& x) { return chain(::atan(x.v), S(1.0 / (1.0 + x.v * x.v)), x); }
KSDD_T KS_DEV DN<N, S> asin(const DN<N, S>& x) { return chain(::asin(x.v), S(1.0 / sqrt(1.0 - x.v * x.v)), x); }
KSDD_T KS_DEV DN<N, S> tan(const DN<N, S>& x) { const S t = ::tan(x.v); return chain(t, S(1.0 + t * t), x); }
KSDD_T KS_DEV DN<N, S> floor(const DN<N, S>& x) { return DN<N, S>(::floor(x.v)); }
KSDD_T KS_DEV DN<N, S> scale2(const DN<N, S>& x, int e) { return x * ldexp(1.0, e); }
KSDD_T KS_DEV DN<N, S> copysign(const DN<N, S>& a, const DN<N, S>& b) { return (signbit(a.v) != signbit(b.v)) ? -a : a; }
KSDD_T KS_DEV DN<N, S>& operator-=(DN<N, S>& a, const DN<N, S>& b) { a = a - b; return a; }
KSDD_T KS_DEV DN<N, S>& operator/=(DN<N, S>& a, const DN<N, S>& b) { a = a / b; return a; }
KSDD_T KS_DEV bool isfinite_(const DN<N, S>& x) { return isfinite_(x.v); }
KSDD_T KS_DEV DN<N, S> atan2(const DN<N, S>& y, const DN<N, S>& x) {
    const S r2 = x.v * x.v + y.v * y.v;
    DN<N, S> r; r.v = atan2(y.v, x.v);
    KSDD_LOOP r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
    return r;
}
// d x^y / dx = y x^y / x: one power per dual power (x^(y - 1) is taken only at x = 0)
KSDD_T KS_DEV DN<N, S> pow(const DN<N, S>& x, double y) {
    const S p = pow(x.v, y);
    return chain(p, y == 0.0 ? S(0.0) : (x.v != 0.0 ? S(y * p / x.v) : S(y * pow(x.v, y - 1.0))), x);
}
KSDD_T KS_DEV DN<N, S> pow(const DN<N, S>& x, const DN<N, S>& y) {
    const S p = pow(x.v, y.v);
    const S dx = y.v == 0.0 ? S(0.0) : (x.v != 0.0 ? S(y.v * p / x.v) : S(y.v * pow(x.v, y.v - 1.0)));
    const S dy = x.v > 0.0 ? S(p * log(x.v)) : S(0.0);
    DN<N, S> r; r.v = p;
    KSDD_LOOP r.d[i] = dx * x.d[i] + dy * y.d[i];
    return r;
}
KSDD_T KS_DEV DN<N, S> pow(double x, const DN<N, S>& y) {
    const S p = pow(x, y.v);
    return chain(p, x > 0.0 ? S(p * log(x)) : S(0.0), y);
}
KSDD_T KS_DEV DN<N, S> dmax(const DN<N, S>& a, const DN<N, S>& b) { return a.v > b.v ? a : b; }
KSDD_T KS_DEV DN<N, S> dmin(const DN<N, S>& a, const DN<N, S>& b) { return a.v < b.v ? a : b; }
KSDD_T KS_DEV DN<N, S> dmax(const DN<N, S>& a, double b) { return a.v > b ? a : DN<N, S>(b); }
KSDD_T KS_DEV DN<N, S> dmin(const DN<N, S>& a, double b) { return a.v < b ? a : DN<N, S>(b); }
#undef KSDD_LOOP
#undef KSDD_T

}  // namespace ksd
