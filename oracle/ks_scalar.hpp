// Scalar types of the CPU oracle.
//
// TEST INFRASTRUCTURE: everything under oracle/ is the FP64 CPU restatement of KerrSplat's hot
// path that the HIP product is checked against. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product never links or calls it.
//
// The oracle is written once, templated on its scalar type `T`:
//   * `double`       - the oracle proper (and the CPU baseline of bench.py);
//   * `Counted`      - a double that counts its own floating-point operations. Running the
//                      oracle on it measures the algorithmic work per ray-sample (the figure the
//                      roofline fraction is computed from): +, -, *, / and sqrt count 1 each, an
//                      fma counts 2, and every transcendental (exp, log, pow, sin, cos, tan, atan,
//                      atan2, asin, acos, tanh, expm1, cbrt, floor) counts 1 and is also tallied
//                      separately as a "special" operation.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <limits>

namespace kso {

struct OpCounts {
    double flops = 0;     // arithmetic operations
    double special = 0;   // transcendental calls (also included in flops)
    void reset() { flops = 0; special = 0; }
};
inline OpCounts& op_counts() { static thread_local OpCounts c; return c; }

struct Counted {
    double v = 0;
    Counted() = default;
    Counted(double x) : v(x) {}
    explicit operator double() const { return v; }
};
inline void cnt(double n = 1) { op_counts().flops += n; }
inline void cnt_special() { op_counts().flops += 1; op_counts().special += 1; }

#define KSO_BINOP(op)                                                                      \
    inline Counted operator op(Counted a, Counted b) { cnt(); return Counted(a.v op b.v); } \
    inline Counted operator op(Counted a, double b) { cnt(); return Counted(a.v op b); }   \
    inline Counted operator op(double a, Counted b) { cnt(); return Counted(a op b.v); }
KSO_BINOP(+)
KSO_BINOP(-)
KSO_BINOP(*)
KSO_BINOP(/)
#undef KSO_BINOP
#define KSO_CMP(op)                                                                   \
    inline bool operator op(Counted a, Counted b) { return a.v op b.v; }              \
    inline bool operator op(Counted a, double b) { return a.v op b; }                 \
    inline bool operator op(double a, Counted b) { return a op b.v; }
KSO_CMP(<)
KSO_CMP(>)
KSO_CMP(<=)
KSO_CMP(>=)
KSO_CMP(==)
KSO_CMP(!=)
#undef KSO_CMP
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted& operator+=(Counted& a, Counted b) { cnt(); a.v += b.v; return a; }
inline Counted& operator-=(Counted& a, Counted b) { cnt(); a.v -= b.v; return a; }
inline Counted& operator*=(Counted& a, Counted b) { cnt(); a.v *= b.v; return a; }
inline Counted& operator/=(Counted& a, Counted b) { cnt(); a.v /= b.v; return a; }

// ---- math functions, overloaded for both scalar types ------------------------------------------
inline double val(double x) { return x; }
inline double val(Counted x) { return x.v; }

#define KSO_UNARY(name, stdname, special)                                        \
    inline double name(double x) { return std::stdname(x); }                     \
    inline Counted name(Counted x) { if (special) cnt_special(); else cnt();     \
                                     return Counted(std::stdname(x.v)); }
KSO_UNARY(ksqrt, sqrt, false)
KSO_UNARY(kexp, exp, true)
KSO_UNARY(klog, log, true)
KSO_UNARY(ksin, sin, true)
KSO_UNARY(kcos, cos, true)
KSO_UNARY(ktan, tan, true)
KSO_UNARY(katan, atan, true)
KSO_UNARY(kasin, asin, true)
KSO_UNARY(kacos, acos, true)
KSO_UNARY(ktanh, tanh, true)
KSO_UNARY(kexpm1, expm1, true)
KSO_UNARY(kcbrt, cbrt, true)
KSO_UNARY(kfloor, floor, true)
KSO_UNARY(ksinh, sinh, true)
KSO_UNARY(kcosh, cosh, true)
#undef KSO_UNARY
inline double kabs(double x) { return std::fabs(x); }
inline Counted kabs(Counted x) { return Counted(std::fabs(x.v)); }
inline double katan2(double y, double x) { return std::atan2(y, x); }
inline Counted katan2(Counted y, Counted x) { cnt_special(); return Counted(std::atan2(y.v, x.v)); }
inline double kpow(double x, double y) { return std::pow(x, y); }
inline Counted kpow(Counted x, Counted y) { cnt_special(); return Counted(std::pow(x.v, y.v)); }
inline Counted kpow(Counted x, double y) { cnt_special(); return Counted(std::pow(x.v, y)); }
inline double kfma(double a, double b, double c) { return std::fma(a, b, c); }
inline Counted kfma(Counted a, Counted b, Counted c) { cnt(2); return Counted(std::fma(a.v, b.v, c.v)); }
inline double kcopysign(double a, double b) { return std::copysign(a, b); }
inline Counted kcopysign(Counted a, Counted b) { return Counted(std::copysign(a.v, b.v)); }
template <class T> inline T kmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T kmin(T a, T b) { return a < b ? a : b; }
template <class T> inline bool kisfinite(T x) { return std::isfinite(val(x)); }
template <class T> inline bool kisnan(T x) { return std::isnan(val(x)); }
template <class T> inline T ksq(T x) { return x * x; }

template <class T> constexpr double PI_D = 3.14159265358979323846;

// A minimal complex number for the quartic roots (std::complex is only defined for the
// built-in floating types).
template <class T> struct Cx {
    T re, im;
    Cx() : re(0.0), im(0.0) {}
    Cx(T r, T i = T(0.0)) : re(r), im(i) {}
};
template <class T> Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> Cx<T> operator-(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> Cx<T> operator-(Cx<T> a) { return {-a.re, -a.im}; }
template <class T> Cx<T> operator*(Cx<T> a, Cx<T> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <class T> Cx<T> operator*(T s, Cx<T> a) { return {s * a.re, s * a.im}; }
template <class T> Cx<T> operator/(Cx<T> a, Cx<T> b) {
    T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <class T> T cabs(Cx<T> a) { return ksqrt(a.re * a.re + a.im * a.im); }
// principal square root
template <class T> Cx<T> csqrt(Cx<T> z) {
    T r = cabs(z);
    if (val(r) == 0.0) return {T(0.0), T(0.0)};
    T s = ksqrt((r + kabs(z.re)) / 2.0);
    if (z.re >= 0.0) return {s, z.im / (2.0 * s)};
    return {kabs(z.im) / (2.0 * s), kcopysign(s, z.im)};
}
// principal cube root
template <class T> Cx<T> ccbrt(Cx<T> z) {
    T r = cabs(z);
    if (val(r) == 0.0) return {T(0.0), T(0.0)};
    T th = katan2(z.im, z.re) / 3.0;
    T m = kcbrt(r);
    return {m * kcos(th), m * ksin(th)};
}

// digamma psi(x) for x > 0: the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 10, then the asymptotic series
// ln x - 1/(2x) - sum B_2k / (2k x^2k) (DLMF 5.5.2, 5.11.2), accurate to rounding
inline double kdigamma(double x) {
    double r = 0.0;
    while (x < 10.0) { r -= 1.0 / x; x += 1.0; }
    const double i2 = 1.0 / (x * x);
    const double ser = i2 * (1.0 / 12.0 - i2 * (1.0 / 120.0 - i2 * (1.0 / 252.0 - i2 * (1.0 / 240.0 - i2 * (1.0 / 132.0 - i2 * (691.0 / 32760.0 - i2 / 12.0))))));
    return r + std::log(x) - 0.5 / x - ser;
}
// the gamma function of a scalar (no operation count); the dual overload in ks_dual.hpp carries Gamma psi
template <class T> T kgamma(const T& x) { return T(std::tgamma(val(x))); }

}  // namespace kso
