// Device elliptic integrals and Jacobi functions (FP64, gfx950).
//
// Carlson's symmetric integrals by duplication (DLMF 19.36.1-2), the incomplete Legendre
// integrals built from them for any real amplitude and any parameter m < 1 (negative parameters
// of the polar motion included), principal values of Pi beyond its pole (DLMF 19.7.9, 19.6.5),
// the Jacobi amplitude by the AGM (DLMF 22.20.1) and the imaginary-modulus transformation for
// m < 0 (DLMF 22.17.2-4). These replace JacobiElliptic.jl's CarlsonAlg.am/K and Krang.jl's
// elliptic integrals (src/Geodesics/jacobi.jl:21-23, quadrature.jl:94 of the reference).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define KS_DEV __host__ __device__ __forceinline__
// Kept out of line: heavy code on rare paths (closed-form anchors, turning points), so that the hot
// loops around the calls keep their registers.
#define KS_COLD __host__ __device__ __attribute__((noinline)) inline

namespace ksd {

constexpr double PI = 3.14159265358979323846;
constexpr double EPS = 2.220446049250313e-16;

KS_DEV double dmax(double a, double b) { return a > b ? a : b; }
KS_DEV double dmin(double a, double b) { return a < b ? a : b; }
KS_DEV double scale2(double x, int e) { return ldexp(x, e); }     // x 2^e

// ---- Carlson ------------------------------------------------------------------------------------
template <class R>
KS_DEV R rf(R x, R y, R z) {
    const R x0 = x, y0 = y;
    R A0 = (x + y + z) * (1.0 / 3.0);
    R Q = 2000.0 * dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z));
    R A = A0, f = 1.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        R sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        R lam = sx * sy + sy * sz + sz * sx;
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25;
        f *= 0.25;
    }
    R Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A;
    R Zd = -(Xd + Yd);
    R E2 = Xd * Yd - Zd * Zd, E3 = Xd * Yd * Zd;
    return (1.0 - E2 * 0.1 + E3 * (1.0 / 14.0) + E2 * E2 * (1.0 / 24.0) - 3.0 * E2 * E3 * (1.0 / 44.0)) / sqrt(A);
}

template <class R>
KS_DEV R rd(R x, R y, R z) {
    const R x0 = x, y0 = y;
    R A0 = (x + y + 3.0 * z) * 0.2;
    R Q = 2000.0 * dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z));
    R A = A0, f = 1.0, sum = 0.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        R sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        R lam = sx * sy + sy * sz + sz * sx;
        sum += f / (sz * (z + lam));
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25;
        f *= 0.25;
    }
    R Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A;
    R Zd = -(Xd + Yd) * (1.0 / 3.0);
    R XY = Xd * Yd, Z2 = Zd * Zd;
    R E2 = XY - 6.0 * Z2, E3 = (3.0 * XY - 8.0 * Z2) * Zd, E4 = 3.0 * (XY - Z2) * Z2, E5 = XY * Z2 * Zd;
    R poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * sqrt(A)) + 3.0 * sum;
}

template <class R>
KS_DEV R rc1(R e) {
    if (fabs(e) < 1e-4) return 1.0 - e * (1.0 / 3.0 - e * (0.2 - e * (1.0 / 7.0 - e * (1.0 / 9.0 - e / 11.0))));
    if (e > 0.0) { R s = sqrt(e); return atan(s) / s; }
    R s = sqrt(-e);
    return 0.5 * log((1.0 + s) / (1.0 - s)) / s;
}

template <class R>
KS_DEV R rj(R x, R y, R z, R p) {
    const R x0 = x, y0 = y, z0 = z;
    R A0 = (x + y + z + 2.0 * p) * 0.2;
    R delta = (p - x) * (p - y) * (p - z);
    R Q = 2000.0 * dmax(dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z)), fabs(A0 - p));
    R A = A0, f = 1.0, sum = 0.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        R sx = sqrt(x), sy = sqrt(y), sz = sqrt(z), sp = sqrt(p);
        R lam = sx * sy + sy * sz + sz * sx;
        R d = (sp + sx) * (sp + sy) * (sp + sz);
        R e = (f * f * f) * delta / (d * d);
        sum += f * rc1<R>(e) / d;
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25; p = (p + lam) * 0.25;
        f *= 0.25;
    }
    R Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A, Zd = f * (A0 - z0) / A;
    R P = -(Xd + Yd + Zd) * 0.5;
    R P2 = P * P;
    R E2 = Xd * Yd + Xd * Zd + Yd * Zd - 3.0 * P2;
    R E3 = Xd * Yd * Zd + 2.0 * E2 * P + 4.0 * P2 * P;
    R E4 = (2.0 * Xd * Yd * Zd + E2 * P + 3.0 * P2 * P) * P;
    R E5 = Xd * Yd * Zd * P2;
    R poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * sqrt(A)) + 6.0 * sum;
}

// ---- Legendre integrals -----------------------------------------------------------------------------
template <class R>
KS_DEV R ellK(R m) { return rf<R>(0.0, 1.0 - m, 1.0); }
template <class R>
KS_DEV R ellEc(R m) { return rf<R>(0.0, 1.0 - m, 1.0) - m * (1.0 / 3.0) * rd<R>(0.0, 1.0 - m, 1.0); }
template <class R>
KS_DEV R ellPic_small(R n, R m) { return rf<R>(0.0, 1.0 - m, 1.0) + n * (1.0 / 3.0) * rj<R>(0.0, 1.0 - m, 1.0, 1.0 - n); }
template <class R>
KS_DEV R ellPic(R n, R m) {
    if (n > 1.0) return ellK<R>(m) - ellPic_small<R>(m / n, m);
    return ellPic_small<R>(n, m);
}

// F, E on |phi| <= pi/2 given (sin, cos) of the amplitude
template <class R>
KS_DEV R ellF_sc(R s, R c, R m) { return s * rf<R>(c * c, 1.0 - m * s * s, 1.0); }
template <class R>
KS_DEV R ellE_sc(R s, R c, R m) {
    R x = c * c, y = 1.0 - m * s * s;
    return s * rf<R>(x, y, 1.0) - m * (1.0 / 3.0) * s * s * s * rd<R>(x, y, 1.0);
}
template <class R>
KS_DEV R ellPi_sc_below(R n, R s, R c, R m) {
    R s2 = s * s;
    R x = c * c, y = 1.0 - m * s2;
    return s * rf<R>(x, y, 1.0) + n * (1.0 / 3.0) * s * s2 * rj<R>(x, y, 1.0, 1.0 - n * s2);
}
template <class R>
KS_DEV R ellPi_sc(R n, R s, R c, R m) {
    R s2 = s * s;
    if (1.0 - n * s2 > 0.0) return ellPi_sc_below<R>(n, s, c, m);
    // principal value beyond the pole, n > 1 (DLMF 19.7.9)
    R pp = sqrt((n - 1.0) * (1.0 - m / n));
    R D = sqrt(1.0 - m * s2);
    R t = s / c;
    return -ellPi_sc_below<R>(m / n, s, c, m) + ellF_sc<R>(s, c, m) + log(fabs((D + pp * t) / (D - pp * t))) / (2.0 * pp);
}

// unfolded to any real amplitude phi = j pi + phi'
template <class R>
KS_DEV R unfold(R phi, R& j) {
    j = floor(phi * (1.0 / PI) + 0.5);
    return phi - j * PI;
}
template <class R>
KS_DEV R ellF(R phi, R m) {
    R j; R p = unfold<R>(phi, j);
    R s, c; sincos(p, &s, &c);
    R r = ellF_sc<R>(s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellK<R>(m) : r;
}
template <class R>
KS_DEV R ellE(R phi, R m) {
    R j; R p = unfold<R>(phi, j);
    R s, c; sincos(p, &s, &c);
    R r = ellE_sc<R>(s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellEc<R>(m) : r;
}
template <class R>
KS_DEV R ellPi(R n, R phi, R m) {
    R j; R p = unfold<R>(phi, j);
    R s, c; sincos(p, &s, &c);
    R r = ellPi_sc<R>(n, s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellPic<R>(n, m) : r;
}

// ---- Jacobi ------------------------------------------------------------------------------------------
// amplitude for 0 <= m < 1 (AGM, descending Landen)
template <class R>
KS_DEV R am(R u, R m) {
    if (m == 0.0) return u;
    R a[10], c[10];
    a[0] = 1.0; R b = sqrt(1.0 - m); c[0] = sqrt(m);
    int N = 0;
    while (N < 9 && fabs(c[N]) > 1e-17 * a[N]) {
        R an = (a[N] + b) * 0.5;
        R bn = sqrt(a[N] * b);
        c[N + 1] = (a[N] - b) * 0.5;
        a[N + 1] = an; b = bn;
        ++N;
    }
    R phi = scale2(a[N] * u, N);
    for (int n = N; n >= 1; --n) phi = (phi + asin(c[n] / a[n] * sin(phi))) * 0.5;
    return phi;
}

template <class R> struct JacT { R sn, cn, dn; };
using Jac = JacT<double>;

// (sn, cn, dn)(u|m), 0 <= m < 1
template <class R>
KS_DEV JacT<R> jac_pos(R u, R m) {
    R ph = am<R>(u, m);
    R s, c; sincos(ph, &s, &c);
    return {s, c, sqrt(1.0 - m * s * s)};
}
// any m < 1
template <class R>
KS_DEV JacT<R> jac(R u, R m) {
    if (m < 0.0) {
        R s = sqrt(1.0 - m), mu = -m / (1.0 - m);
        JacT<R> x = jac_pos<R>(s * u, mu);
        return {x.sn / (s * x.dn), x.cn / x.dn, 1.0 / x.dn};
    }
    return jac_pos<R>(u, m);
}
// continuous amplitude for any m < 1
template <class R>
KS_DEV R am_any(R u, R m) {
    if (!(m < 0.0)) return am<R>(u, m);
    R s = sqrt(1.0 - m), mu = -m / (1.0 - m);
    R K2 = 2.0 * ellK<R>(mu) / s;
    R j = floor(u / K2 + 0.5);
    R ur = u - j * K2;
    R ph = am<R>(s * ur, mu);
    R sn, cn; sincos(ph, &sn, &cn);
    return atan2(sn / s, cn) + j * PI;
}

}  // namespace ksd
