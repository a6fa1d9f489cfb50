// Forward-mode dual numbers for the oracle's adjoint (test infrastructure; see ks_scalar.hpp).
// Dual<N> carries a value and N partials; the oracle's templated functions (step operator,
// coefficients, frames, splat weight) run on it unchanged, which is how the reference computes the
// per-sample local derivatives of its dual sweep (ForwardDiff duals, dual_gradient.jl:1-19).
#pragma once
#include "ks_scalar.hpp"

namespace kso {

template <int N> struct Dual {
    double v = 0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
    static Dual seed(double x, int i) { Dual r(x); r.d[i] = 1.0; return r; }
};
template <int N> double val(const Dual<N>& x) { return x.v; }

#define KSD_LOOP for (int i = 0; i < N; ++i)
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v + b.v); KSD_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v - b.v); KSD_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v * b.v); KSD_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v); double ib = 1.0 / b.v;
    KSD_LOOP r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { Dual<N> r = b; r.v += a; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { Dual<N> r(a - b.v); KSD_LOOP r.d[i] = -b.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { Dual<N> r(a.v * b); KSD_LOOP r.d[i] = a.d[i] * b; return r; }
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { return a * (1.0 / b); }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { Dual<N> r(a / b.v); double f = -r.v / b.v; KSD_LOOP r.d[i] = f * b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a) { Dual<N> r(-a.v); KSD_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
template <int N> Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { a = a * b; return a; }
#define KSD_CMP(op)                                                                                  \
    template <int N> bool operator op(const Dual<N>& a, const Dual<N>& b) { return a.v op b.v; }    \
    template <int N> bool operator op(const Dual<N>& a, double b) { return a.v op b; }              \
    template <int N> bool operator op(double a, const Dual<N>& b) { return a op b.v; }
KSD_CMP(<)
KSD_CMP(>)
KSD_CMP(<=)
KSD_CMP(>=)
KSD_CMP(==)
KSD_CMP(!=)
#undef KSD_CMP

// chain rule helper
template <int N> Dual<N> chain(double f, double df, const Dual<N>& x) { Dual<N> r(f); KSD_LOOP r.d[i] = df * x.d[i]; return r; }
template <int N> Dual<N> ksqrt(const Dual<N>& x) { double s = std::sqrt(x.v); return chain(s, s > 0 ? 0.5 / s : 0.0, x); }
template <int N> Dual<N> kexp(const Dual<N>& x) { double e = std::exp(x.v); return chain(e, e, x); }
template <int N> Dual<N> kexpm1(const Dual<N>& x) { return chain(std::expm1(x.v), std::exp(x.v), x); }
template <int N> Dual<N> klog(const Dual<N>& x) { return chain(std::log(x.v), 1.0 / x.v, x); }
template <int N> Dual<N> ksin(const Dual<N>& x) { return chain(std::sin(x.v), std::cos(x.v), x); }
template <int N> Dual<N> kcos(const Dual<N>& x) { return chain(std::cos(x.v), -std::sin(x.v), x); }
template <int N> Dual<N> ktan(const Dual<N>& x) { double t = std::tan(x.v); return chain(t, 1.0 + t * t, x); }
template <int N> Dual<N> katan(const Dual<N>& x) { return chain(std::atan(x.v), 1.0 / (1.0 + x.v * x.v), x); }
template <int N> Dual<N> kasin(const Dual<N>& x) { return chain(std::asin(x.v), 1.0 / std::sqrt(1.0 - x.v * x.v), x); }
// acos with a zero derivative at the ends (the reference's safe_acos, step.jl:285-291)
template <int N> Dual<N> kacos(const Dual<N>& x) {
    double v = x.v, s = 1.0 - v * v;
    return chain(std::acos(v), s > 0 ? -1.0 / std::sqrt(s) : 0.0, x);
}
template <int N> Dual<N> ktanh(const Dual<N>& x) { double t = std::tanh(x.v); return chain(t, 1.0 - t * t, x); }
template <int N> Dual<N> ksinh(const Dual<N>& x) { return chain(std::sinh(x.v), std::cosh(x.v), x); }
template <int N> Dual<N> kcosh(const Dual<N>& x) { return chain(std::cosh(x.v), std::sinh(x.v), x); }
template <int N> Dual<N> kcbrt(const Dual<N>& x) { double c = std::cbrt(x.v); return chain(c, c / (3.0 * x.v), x); }
template <int N> Dual<N> kfloor(const Dual<N>& x) { return Dual<N>(std::floor(x.v)); }
template <int N> Dual<N> kabs(const Dual<N>& x) { return x.v < 0 ? -x : x; }
template <int N> Dual<N> kcopysign(const Dual<N>& a, const Dual<N>& b) { return (std::signbit(a.v) != std::signbit(b.v)) ? -a : a; }
template <int N> Dual<N> katan2(const Dual<N>& y, const Dual<N>& x) {
    double r2 = x.v * x.v + y.v * y.v;
    Dual<N> r(std::atan2(y.v, x.v));
    KSD_LOOP r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
    return r;
}
template <int N> Dual<N> kpow(const Dual<N>& x, double y) {
    double p = std::pow(x.v, y);
    return chain(p, y == 0.0 ? 0.0 : y * std::pow(x.v, y - 1.0), x);
}
template <int N> Dual<N> kpow(const Dual<N>& x, const Dual<N>& y) {
    double p = std::pow(x.v, y.v);
    Dual<N> r(p);
    double lx = x.v > 0 ? std::log(x.v) : 0.0;
    double dx = y.v == 0.0 ? 0.0 : y.v * std::pow(x.v, y.v - 1.0);
    KSD_LOOP r.d[i] = dx * x.d[i] + p * lx * y.d[i];
    return r;
}
#undef KSD_LOOP

}  // namespace kso
