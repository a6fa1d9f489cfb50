// Elliptic integrals and Jacobi elliptic functions for the CPU oracle (test infrastructure; see
// ks_scalar.hpp). Parameter convention: m = k^2 (DLMF 19.2, 22.2).
//
// * Carlson's symmetric integrals R_F, R_D, R_J by the duplication algorithm (DLMF 19.36.1-2,
//   Carlson 1995), and from them the incomplete Legendre integrals F, E, Pi (DLMF 19.25.5,
//   19.25.7, 19.25.14) for any real amplitude (unfolded with the complete integrals) and any
//   parameter m < 1, including the negative parameters of the polar motion.
// * The principal value of Pi(n; phi|m) beyond its pole (n > 1, n sin^2 phi > 1) through DLMF
//   19.7.9 and 19.6.5: Pi(n) = -Pi(m/n) + F + ln|(D + p tan)/(D - p tan)|/(2p).
// * The Jacobi amplitude by the arithmetic-geometric mean (DLMF 22.20.1, A&S 16.4), and for a
//   negative parameter through the imaginary-modulus transformation (DLMF 22.17.2-4).
//
// These are the functions the reference takes from JacobiElliptic.jl (CarlsonAlg.am, K) and
// Krang.jl's elliptic integrals (src/Geodesics/jacobi.jl:21-23, quadrature.jl:94).
#pragma once
#include "ks_scalar.hpp"

namespace kso {

constexpr double KS_PI = 3.14159265358979323846;
constexpr double KS_EPS = 2.220446049250313e-16;

// ---- Carlson's symmetric forms -------------------------------------------------------------
template <class T> T carlson_rf(T x, T y, T z) {
    T A0 = (x + y + z) / 3.0;
    T Q = 2000.0 * kmax(kmax(kabs(A0 - x), kabs(A0 - y)), kabs(A0 - z));   // (3 eps)^(-1/6) ~ 1.8e2, with margin
    T A = A0, xn = x, yn = y, zn = z;
    double f = 1.0;
    for (int it = 0; it < 100 && val(Q) * f >= val(kabs(A)); ++it) {
        T sx = ksqrt(xn), sy = ksqrt(yn), sz = ksqrt(zn);
        T lam = sx * sy + sy * sz + sz * sx;
        A = (A + lam) / 4.0; xn = (xn + lam) / 4.0; yn = (yn + lam) / 4.0; zn = (zn + lam) / 4.0;
        f *= 0.25;
    }
    T X = f * (A0 - x) / A, Y = f * (A0 - y) / A;
    T Z = -(X + Y);
    T E2 = X * Y - Z * Z, E3 = X * Y * Z;
    return (1.0 - E2 / 10.0 + E3 / 14.0 + E2 * E2 / 24.0 - 3.0 * E2 * E3 / 44.0) / ksqrt(A);
}

template <class T> T carlson_rd(T x, T y, T z) {
    T A0 = (x + y + 3.0 * z) / 5.0;
    T Q = 2000.0 * kmax(kmax(kabs(A0 - x), kabs(A0 - y)), kabs(A0 - z));
    T A = A0, xn = x, yn = y, zn = z, sum = 0.0;
    double f = 1.0;
    for (int it = 0; it < 100 && val(Q) * f >= val(kabs(A)); ++it) {
        T sx = ksqrt(xn), sy = ksqrt(yn), sz = ksqrt(zn);
        T lam = sx * sy + sy * sz + sz * sx;
        sum = sum + f / (sz * (zn + lam));
        A = (A + lam) / 4.0; xn = (xn + lam) / 4.0; yn = (yn + lam) / 4.0; zn = (zn + lam) / 4.0;
        f *= 0.25;
    }
    T X = f * (A0 - x) / A, Y = f * (A0 - y) / A;
    T Z = -(X + Y) / 3.0;
    T XY = X * Y, Z2 = Z * Z;
    T E2 = XY - 6.0 * Z2, E3 = (3.0 * XY - 8.0 * Z2) * Z, E4 = 3.0 * (XY - Z2) * Z2, E5 = XY * Z2 * Z;
    T poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * ksqrt(A)) + 3.0 * sum;
}

// R_C(1, 1 + e) for e > -1 (the degenerate integral inside R_J's duplication)
template <class T> T carlson_rc1(T e) {
    if (val(kabs(e)) < 1e-4) {
        // series: 1 - e/3 + e^2/5 - e^3/7 + e^4/9 - e^5/11
        return 1.0 - e * (1.0 / 3.0 - e * (1.0 / 5.0 - e * (1.0 / 7.0 - e * (1.0 / 9.0 - e / 11.0))));
    }
    if (e > 0.0) { T s = ksqrt(e); return katan(s) / s; }
    T s = ksqrt(-e);
    return 0.5 * klog((1.0 + s) / (1.0 - s)) / s;     // atanh(s)/s
}

template <class T> T carlson_rj(T x, T y, T z, T p) {
    T A0 = (x + y + z + 2.0 * p) / 5.0;
    T delta = (p - x) * (p - y) * (p - z);
    T Q = 2000.0 * kmax(kmax(kmax(kabs(A0 - x), kabs(A0 - y)), kabs(A0 - z)), kabs(A0 - p));
    T A = A0, xn = x, yn = y, zn = z, pn = p, sum = 0.0;
    double f = 1.0;
    for (int it = 0; it < 100 && val(Q) * f >= val(kabs(A)); ++it) {
        T sx = ksqrt(xn), sy = ksqrt(yn), sz = ksqrt(zn), sp = ksqrt(pn);
        T lam = sx * sy + sy * sz + sz * sx;
        T d = (sp + sx) * (sp + sy) * (sp + sz);
        T e = (f * f * f) * delta / (d * d);
        sum = sum + f * carlson_rc1(e) / d;
        A = (A + lam) / 4.0; xn = (xn + lam) / 4.0; yn = (yn + lam) / 4.0; zn = (zn + lam) / 4.0; pn = (pn + lam) / 4.0;
        f *= 0.25;
    }
    T X = f * (A0 - x) / A, Y = f * (A0 - y) / A, Z = f * (A0 - z) / A;
    T P = -(X + Y + Z) / 2.0;
    T P2 = P * P;
    T E2 = X * Y + X * Z + Y * Z - 3.0 * P2;
    T E3 = X * Y * Z + 2.0 * E2 * P + 4.0 * P2 * P;
    T E4 = (2.0 * X * Y * Z + E2 * P + 3.0 * P2 * P) * P;
    T E5 = X * Y * Z * P2;
    T poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * ksqrt(A)) + 6.0 * sum;
}

// ---- complete integrals ----------------------------------------------------------------------
template <class T> T ellK(T m) { return carlson_rf(T(0.0), 1.0 - m, T(1.0)); }
template <class T> T ellE(T m) { return carlson_rf(T(0.0), 1.0 - m, T(1.0)) - m / 3.0 * carlson_rd(T(0.0), 1.0 - m, T(1.0)); }
// complete Pi(n|m); principal value for n > 1 (DLMF 19.6.5)
template <class T> T ellPi_c(T n, T m) {
    if (n > 1.0) return ellK(m) - ellPi_c(m / n, m);
    return carlson_rf(T(0.0), 1.0 - m, T(1.0)) + n / 3.0 * carlson_rj(T(0.0), 1.0 - m, T(1.0), 1.0 - n);
}

// ---- incomplete integrals on the principal amplitude range |phi| <= pi/2 ---------------------
template <class T> T ellF_p(T phi, T m) {
    T s = ksin(phi), c = kcos(phi);
    return s * carlson_rf(c * c, 1.0 - m * s * s, T(1.0));
}
template <class T> T ellE_p(T phi, T m) {
    T s = ksin(phi), c = kcos(phi);
    T x = c * c, y = 1.0 - m * s * s;
    return s * carlson_rf(x, y, T(1.0)) - m / 3.0 * s * s * s * carlson_rd(x, y, T(1.0));
}
template <class T> T ellPi_p(T n, T phi, T m) {
    T s = ksin(phi), c = kcos(phi);
    T s2 = s * s;
    T p = 1.0 - n * s2;
    if (p > 0.0) {
        T x = c * c, y = 1.0 - m * s2;
        return s * carlson_rf(x, y, T(1.0)) + n / 3.0 * s * s2 * carlson_rj(x, y, T(1.0), p);
    }
    // principal value beyond the pole (n > 1): DLMF 19.7.9
    T pp = ksqrt((n - 1.0) * (1.0 - m / n));
    T D = ksqrt(1.0 - m * s2);
    T t = s / c;
    return -ellPi_p(m / n, phi, m) + ellF_p(phi, m) + klog(kabs((D + pp * t) / (D - pp * t))) / (2.0 * pp);
}

// ---- unfolded to any real amplitude --------------------------------------------------------------
// phi = j pi + phi', |phi'| <= pi/2: F(phi) = 2 j K + F(phi'), and likewise for E and Pi.
template <class T> inline T unfold(T phi, double& j) {
    j = std::floor(val(phi) / KS_PI + 0.5);
    return phi - j * KS_PI;
}
template <class T> T ellF(T phi, T m) {
    double j; T p = unfold(phi, j);
    T r = ellF_p(p, m);
    return j != 0.0 ? r + 2.0 * j * ellK(m) : r;
}
template <class T> T ellE(T phi, T m) {
    double j; T p = unfold(phi, j);
    T r = ellE_p(p, m);
    return j != 0.0 ? r + 2.0 * j * ellE(m) : r;
}
template <class T> T ellPi(T n, T phi, T m) {
    double j; T p = unfold(phi, j);
    T r = ellPi_p(n, p, m);
    return j != 0.0 ? r + 2.0 * j * ellPi_c(n, m) : r;
}

// ---- Jacobi functions ------------------------------------------------------------------------------
template <class T> struct Jac { T sn, cn, dn; };

// amplitude am(u|m) for 0 <= m < 1 by the AGM (DLMF 22.20.1)
template <class T> T jacobi_am(T u, T m) {
    if (val(m) == 0.0) return u;
    T a[16], c[16];
    a[0] = 1.0; T b = ksqrt(1.0 - m); c[0] = ksqrt(m);
    int N = 0;
    while (N < 15 && val(kabs(c[N])) > 1e-17 * val(a[N])) {
        T an = (a[N] + b) / 2.0;
        T bn = ksqrt(a[N] * b);
        c[N + 1] = (a[N] - b) / 2.0;
        a[N + 1] = an; b = bn;
        ++N;
    }
    T phi = std::ldexp(1.0, N) * a[N] * u;
    for (int n = N; n >= 1; --n) phi = (phi + kasin(c[n] / a[n] * ksin(phi))) / 2.0;
    return phi;
}

// (sn, cn, dn)(u|m) for any m < 1; negative m through the imaginary-modulus transformation
// sn(u|m) = sd(su|mu)/s, cn(u|m) = cd(su|mu), dn(u|m) = nd(su|mu), s = sqrt(1-m), mu = -m/(1-m).
template <class T> Jac<T> jacobi(T u, T m) {
    if (m < 0.0) {
        T s = ksqrt(1.0 - m), mu = -m / (1.0 - m);
        Jac<T> x = jacobi(s * u, mu);
        return {x.sn / (s * x.dn), x.cn / x.dn, 1.0 / x.dn};
    }
    T phi = jacobi_am(u, m);
    T sn = ksin(phi), cn = kcos(phi);
    return {sn, cn, ksqrt(1.0 - m * sn * sn)};
}

// continuous amplitude for any m < 1 (negative m: atan2 of the transformed functions, unfolded
// by the period 2K(m) = 2K(mu)/s)
template <class T> T jacobi_am_any(T u, T m) {
    if (!(m < 0.0)) return jacobi_am(u, m);
    T s = ksqrt(1.0 - m), mu = -m / (1.0 - m);
    T K2 = 2.0 * ellK(mu) / s;
    double j = std::floor(val(u / K2) + 0.5);
    T ur = u - j * K2;
    T ph = jacobi_am(s * ur, mu);
    T sn = ksin(ph), cn = kcos(ph);
    return katan2(sn / s, cn) + j * KS_PI;
}

}  // namespace kso
