// Forward-mode dual numbers for the device adjoint: DN<N> carries a value and N partials. The
// transfer functions are templates on their scalar type, so the same source gives the FP64 forward
// (double) and the per-sample local derivatives of the adjoint (DN<7>), as the reference obtains them
// with ForwardDiff duals (src/Splats/dual_gradient.jl:1-19, src/Transfer/transport.jl:332-341).
#pragma once
#include "ks_dmath.hpp"

namespace ksd {

// the double overloads stay visible beside the dual ones declared in this namespace
using ::sqrt; using ::exp; using ::expm1; using ::log; using ::log1p; using ::sin; using ::cos; using ::sincos;
using ::sinh; using ::cosh; using ::tanh; using ::cbrt; using ::acos; using ::fabs; using ::atan2; using ::pow;

template <int N> struct DN {
    double v;
    double d[N];
    KS_DEV DN() {}
    KS_DEV DN(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    KS_DEV static DN seed(double x, int k) { DN r(x); r.d[k] = 1.0; return r; }
};

// the value of a scalar, dual or not
KS_DEV double val(double x) { return x; }
template <int N> KS_DEV double val(const DN<N>& x) { return x.v; }

#define KSDD_LOOP _Pragma("unroll") for (int i = 0; i < N; ++i)
template <int N> KS_DEV DN<N> operator+(const DN<N>& a, const DN<N>& b) { DN<N> r; r.v = a.v + b.v; KSDD_LOOP r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> KS_DEV DN<N> operator-(const DN<N>& a, const DN<N>& b) { DN<N> r; r.v = a.v - b.v; KSDD_LOOP r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> KS_DEV DN<N> operator*(const DN<N>& a, const DN<N>& b) { DN<N> r; r.v = a.v * b.v; KSDD_LOOP r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> KS_DEV DN<N> operator/(const DN<N>& a, const DN<N>& b) {
    DN<N> r; r.v = a.v / b.v; const double ib = 1.0 / b.v;
    KSDD_LOOP r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> KS_DEV DN<N> operator+(const DN<N>& a, double b) { DN<N> r = a; r.v += b; return r; }
template <int N> KS_DEV DN<N> operator+(double a, const DN<N>& b) { DN<N> r = b; r.v += a; return r; }
template <int N> KS_DEV DN<N> operator-(const DN<N>& a, double b) { DN<N> r = a; r.v -= b; return r; }
template <int N> KS_DEV DN<N> operator-(double a, const DN<N>& b) { DN<N> r; r.v = a - b.v; KSDD_LOOP r.d[i] = -b.d[i]; return r; }
template <int N> KS_DEV DN<N> operator*(const DN<N>& a, double b) { DN<N> r; r.v = a.v * b; KSDD_LOOP r.d[i] = a.d[i] * b; return r; }
template <int N> KS_DEV DN<N> operator*(double a, const DN<N>& b) { return b * a; }
template <int N> KS_DEV DN<N> operator/(const DN<N>& a, double b) { return a * (1.0 / b); }
template <int N> KS_DEV DN<N> operator/(double a, const DN<N>& b) { DN<N> r; r.v = a / b.v; const double f = -r.v / b.v; KSDD_LOOP r.d[i] = f * b.d[i]; return r; }
template <int N> KS_DEV DN<N> operator-(const DN<N>& a) { DN<N> r; r.v = -a.v; KSDD_LOOP r.d[i] = -a.d[i]; return r; }
template <int N> KS_DEV DN<N>& operator+=(DN<N>& a, const DN<N>& b) { a = a + b; return a; }
template <int N> KS_DEV DN<N>& operator*=(DN<N>& a, const DN<N>& b) { a = a * b; return a; }
template <int N> KS_DEV DN<N>& operator*=(DN<N>& a, double b) { a = a * b; return a; }
#define KSDD_CMP(op)                                                                                      \
    template <int N> KS_DEV bool operator op(const DN<N>& a, const DN<N>& b) { return a.v op b.v; }      \
    template <int N> KS_DEV bool operator op(const DN<N>& a, double b) { return a.v op b; }              \
    template <int N> KS_DEV bool operator op(double a, const DN<N>& b) { return a op b.v; }
KSDD_CMP(<)
KSDD_CMP(>)
KSDD_CMP(<=)
KSDD_CMP(>=)
KSDD_CMP(==)
KSDD_CMP(!=)
#undef KSDD_CMP

template <int N> KS_DEV DN<N> chain(double f, double df, const DN<N>& x) { DN<N> r; r.v = f; KSDD_LOOP r.d[i] = df * x.d[i]; return r; }
template <int N> KS_DEV DN<N> sqrt(const DN<N>& x) { const double s = ::sqrt(x.v); return chain(s, s > 0.0 ? 0.5 / s : 0.0, x); }
template <int N> KS_DEV DN<N> exp(const DN<N>& x) { const double e = ::exp(x.v); return chain(e, e, x); }
template <int N> KS_DEV DN<N> expm1(const DN<N>& x) { return chain(::expm1(x.v), ::exp(x.v), x); }
template <int N> KS_DEV DN<N> log(const DN<N>& x) { return chain(::log(x.v), 1.0 / x.v, x); }
template <int N> KS_DEV DN<N> log1p(const DN<N>& x) { return chain(::log1p(x.v), 1.0 / (1.0 + x.v), x); }
template <int N> KS_DEV DN<N> sin(const DN<N>& x) { double s, c; ::sincos(x.v, &s, &c); return chain(s, c, x); }
template <int N> KS_DEV DN<N> cos(const DN<N>& x) { double s, c; ::sincos(x.v, &s, &c); return chain(c, -s, x); }
template <int N> KS_DEV void sincos(const DN<N>& x, DN<N>* s, DN<N>* c) {
    double sv, cv; ::sincos(x.v, &sv, &cv);
    *s = chain(sv, cv, x); *c = chain(cv, -sv, x);
}
template <int N> KS_DEV DN<N> sinh(const DN<N>& x) { return chain(::sinh(x.v), ::cosh(x.v), x); }
template <int N> KS_DEV DN<N> cosh(const DN<N>& x) { return chain(::cosh(x.v), ::sinh(x.v), x); }
template <int N> KS_DEV DN<N> tanh(const DN<N>& x) { const double t = ::tanh(x.v); return chain(t, 1.0 - t * t, x); }
template <int N> KS_DEV DN<N> cbrt(const DN<N>& x) { const double c = ::cbrt(x.v); return chain(c, c / (3.0 * x.v), x); }
// acos with a zero derivative at the ends (the reference's safe_acos, step.jl:285-291)
template <int N> KS_DEV DN<N> acos(const DN<N>& x) {
    const double s = 1.0 - x.v * x.v;
    return chain(::acos(x.v), s > 0.0 ? -1.0 / ::sqrt(s) : 0.0, x);
}
template <int N> KS_DEV DN<N> fabs(const DN<N>& x) { return x.v < 0.0 ? -x : x; }
template <int N> KS_DEV DN<N> atan2(const DN<N>& y, const DN<N>& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    DN<N> r; r.v = ::atan2(y.v, x.v);
    KSDD_LOOP r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
    return r;
}
template <int N> KS_DEV DN<N> pow(const DN<N>& x, double y) {
    const double p = ::pow(x.v, y);
    return chain(p, y == 0.0 ? 0.0 : y * ::pow(x.v, y - 1.0), x);
}
template <int N> KS_DEV DN<N> dmax(const DN<N>& a, const DN<N>& b) { return a.v > b.v ? a : b; }
template <int N> KS_DEV DN<N> dmin(const DN<N>& a, const DN<N>& b) { return a.v < b.v ? a : b; }
template <int N> KS_DEV DN<N> dmax(const DN<N>& a, double b) { return a.v > b ? a : DN<N>(b); }
template <int N> KS_DEV DN<N> dmin(const DN<N>& a, double b) { return a.v < b ? a : DN<N>(b); }
#undef KSDD_LOOP

}  // namespace ksd
