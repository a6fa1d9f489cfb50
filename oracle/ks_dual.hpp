// Forward-mode dual numbers for the oracle's adjoint (test infrastructure; see ks_scalar.hpp).
// Dual<N, S> carries a value and N partials of the base scalar S (double, or Counted to measure the
// adjoint's algorithmic work); the oracle's templated functions (step operator, coefficients, frames,
// splat weight) run on it unchanged, which is how the reference computes the per-sample local
// derivatives of its dual sweep (ForwardDiff duals, dual_gradient.jl:1-19).
#pragma once
#include <type_traits>
#include "ks_scalar.hpp"

namespace kso {

template <class S> struct ident { using type = S; };

template <int N, class S = double> struct Dual {
    S v;
    S d[N];
    Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = S(0.0); }
    Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = S(0.0); }
    template <class X, class = std::enable_if_t<std::is_same_v<X, S> && !std::is_same_v<S, double>>>
    Dual(X x) : v(x) { for (int i = 0; i < N; ++i) d[i] = S(0.0); }
    static Dual seed(S x, int i) { Dual r; r.v = x; r.d[i] = S(1.0); return r; }
};
template <int N, class S> double val(const Dual<N, S>& x) { return val(x.v); }

#define KSD_LOOP for (int i = 0; i < N; ++i)
#define KSD_T template <int N, class S>
KSD_T Dual<N, S> operator+(const Dual<N, S>& a, const Dual<N, S>& b) { Dual<N, S> r; r.v = a.v + b.v; KSD_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
KSD_T Dual<N, S> operator-(const Dual<N, S>& a, const Dual<N, S>& b) { Dual<N, S> r; r.v = a.v - b.v; KSD_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
KSD_T Dual<N, S> operator*(const Dual<N, S>& a, const Dual<N, S>& b) { Dual<N, S> r; r.v = a.v * b.v; KSD_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
KSD_T Dual<N, S> operator/(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r; r.v = a.v / b.v; S ib = 1.0 / b.v;
    KSD_LOOP r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
KSD_T Dual<N, S> operator+(const Dual<N, S>& a, double b) { Dual<N, S> r = a; r.v = r.v + b; return r; }
KSD_T Dual<N, S> operator+(double a, const Dual<N, S>& b) { Dual<N, S> r = b; r.v = r.v + a; return r; }
KSD_T Dual<N, S> operator-(const Dual<N, S>& a, double b) { Dual<N, S> r = a; r.v = r.v - b; return r; }
KSD_T Dual<N, S> operator-(double a, const Dual<N, S>& b) { Dual<N, S> r; r.v = a - b.v; KSD_LOOP r.d[i] = -b.d[i]; return r; }
KSD_T Dual<N, S> operator*(const Dual<N, S>& a, double b) { Dual<N, S> r; r.v = a.v * b; KSD_LOOP r.d[i] = a.d[i] * b; return r; }
KSD_T Dual<N, S> operator*(double a, const Dual<N, S>& b) { return b * a; }
KSD_T Dual<N, S> operator/(const Dual<N, S>& a, double b) { return a * (1.0 / b); }
KSD_T Dual<N, S> operator/(double a, const Dual<N, S>& b) { Dual<N, S> r; r.v = a / b.v; S f = -r.v / b.v; KSD_LOOP r.d[i] = f * b.d[i]; return r; }
KSD_T Dual<N, S> operator-(const Dual<N, S>& a) { Dual<N, S> r; r.v = -a.v; KSD_LOOP r.d[i] = -a.d[i]; return r; }
KSD_T Dual<N, S>& operator+=(Dual<N, S>& a, const Dual<N, S>& b) { a = a + b; return a; }
KSD_T Dual<N, S>& operator-=(Dual<N, S>& a, const Dual<N, S>& b) { a = a - b; return a; }
KSD_T Dual<N, S>& operator*=(Dual<N, S>& a, const Dual<N, S>& b) { a = a * b; return a; }
#define KSD_CMP(op)                                                                              \
    KSD_T bool operator op(const Dual<N, S>& a, const Dual<N, S>& b) { return val(a.v) op val(b.v); } \
    KSD_T bool operator op(const Dual<N, S>& a, double b) { return val(a.v) op b; }                   \
    KSD_T bool operator op(double a, const Dual<N, S>& b) { return a op val(b.v); }
KSD_CMP(<)
KSD_CMP(>)
KSD_CMP(<=)
KSD_CMP(>=)
KSD_CMP(==)
KSD_CMP(!=)
#undef KSD_CMP

// chain rule helper
KSD_T Dual<N, S> chain(typename ident<S>::type f, typename ident<S>::type df, const Dual<N, S>& x) {
    Dual<N, S> r; r.v = f; KSD_LOOP r.d[i] = df * x.d[i]; return r;
}
KSD_T Dual<N, S> ksqrt(const Dual<N, S>& x) { S s = ksqrt(x.v); return chain(s, s > 0.0 ? S(0.5 / s) : S(0.0), x); }
KSD_T Dual<N, S> kexp(const Dual<N, S>& x) { S e = kexp(x.v); return chain(e, e, x); }
KSD_T Dual<N, S> kexpm1(const Dual<N, S>& x) { return chain(kexpm1(x.v), kexp(x.v), x); }
KSD_T Dual<N, S> klog(const Dual<N, S>& x) { return chain(klog(x.v), S(1.0 / x.v), x); }
KSD_T Dual<N, S> ksin(const Dual<N, S>& x) { return chain(ksin(x.v), kcos(x.v), x); }
KSD_T Dual<N, S> kcos(const Dual<N, S>& x) { return chain(kcos(x.v), S(-ksin(x.v)), x); }
KSD_T Dual<N, S> ktan(const Dual<N, S>& x) { S t = ktan(x.v); return chain(t, S(1.0 + t * t), x); }
KSD_T Dual<N, S> katan(const Dual<N, S>& x) { return chain(katan(x.v), S(1.0 / (1.0 + x.v * x.v)), x); }
KSD_T Dual<N, S> kasin(const Dual<N, S>& x) { return chain(kasin(x.v), S(1.0 / ksqrt(1.0 - x.v * x.v)), x); }
// acos with a zero derivative at the ends (the reference's safe_acos, step.jl:285-291)
KSD_T Dual<N, S> kacos(const Dual<N, S>& x) {
    S s = 1.0 - x.v * x.v;
    return chain(kacos(x.v), s > 0.0 ? S(-1.0 / ksqrt(s)) : S(0.0), x);
}
KSD_T Dual<N, S> ktanh(const Dual<N, S>& x) { S t = ktanh(x.v); return chain(t, S(1.0 - t * t), x); }
KSD_T Dual<N, S> ksinh(const Dual<N, S>& x) { return chain(ksinh(x.v), kcosh(x.v), x); }
KSD_T Dual<N, S> kcosh(const Dual<N, S>& x) { return chain(kcosh(x.v), ksinh(x.v), x); }
KSD_T Dual<N, S> kcbrt(const Dual<N, S>& x) { S c = kcbrt(x.v); return chain(c, S(c / (3.0 * x.v)), x); }
KSD_T Dual<N, S> kgamma(const Dual<N, S>& x) { const double g = std::tgamma(val(x.v)); return chain(S(g), S(g * kdigamma(val(x.v))), x); }
KSD_T Dual<N, S> kfloor(const Dual<N, S>& x) { return Dual<N, S>(kfloor(x.v)); }
KSD_T Dual<N, S> kabs(const Dual<N, S>& x) { return val(x.v) < 0 ? -x : x; }
KSD_T Dual<N, S> kcopysign(const Dual<N, S>& a, const Dual<N, S>& b) { return (std::signbit(val(a.v)) != std::signbit(val(b.v))) ? -a : a; }
KSD_T Dual<N, S> katan2(const Dual<N, S>& y, const Dual<N, S>& x) {
    S r2 = x.v * x.v + y.v * y.v;
    Dual<N, S> r; r.v = katan2(y.v, x.v);
    KSD_LOOP r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
    return r;
}
KSD_T Dual<N, S> kpow(const Dual<N, S>& x, double y) {
    S p = kpow(x.v, y);
    return chain(p, y == 0.0 ? S(0.0) : S(y * kpow(x.v, y - 1.0)), x);
}
KSD_T Dual<N, S> kpow(const Dual<N, S>& x, const Dual<N, S>& y) {
    S p = kpow(x.v, y.v);
    Dual<N, S> r; r.v = p;
    S lx = x.v > 0.0 ? S(klog(x.v)) : S(0.0);
    S dx = y.v == 0.0 ? S(0.0) : S(y.v * kpow(x.v, y.v - 1.0));
    KSD_LOOP r.d[i] = dx * x.d[i] + p * lx * y.d[i];
    return r;
}
#undef KSD_T
#undef KSD_LOOP

}  // namespace kso
