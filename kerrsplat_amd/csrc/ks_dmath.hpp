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

// ---- Carlson ------------------------------------------------------------------------------------
KS_DEV double rf(double x, double y, double z) {
    const double x0 = x, y0 = y;
    double A0 = (x + y + z) * (1.0 / 3.0);
    double Q = 2000.0 * dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z));
    double A = A0, f = 1.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        double lam = sx * sy + sy * sz + sz * sx;
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25;
        f *= 0.25;
    }
    double Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A;
    double Zd = -(Xd + Yd);
    double E2 = Xd * Yd - Zd * Zd, E3 = Xd * Yd * Zd;
    return (1.0 - E2 * 0.1 + E3 * (1.0 / 14.0) + E2 * E2 * (1.0 / 24.0) - 3.0 * E2 * E3 * (1.0 / 44.0)) / sqrt(A);
}

KS_DEV double rd(double x, double y, double z) {
    const double x0 = x, y0 = y;
    double A0 = (x + y + 3.0 * z) * 0.2;
    double Q = 2000.0 * dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z));
    double A = A0, f = 1.0, sum = 0.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        double lam = sx * sy + sy * sz + sz * sx;
        sum += f / (sz * (z + lam));
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25;
        f *= 0.25;
    }
    double Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A;
    double Zd = -(Xd + Yd) * (1.0 / 3.0);
    double XY = Xd * Yd, Z2 = Zd * Zd;
    double E2 = XY - 6.0 * Z2, E3 = (3.0 * XY - 8.0 * Z2) * Zd, E4 = 3.0 * (XY - Z2) * Z2, E5 = XY * Z2 * Zd;
    double poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * sqrt(A)) + 3.0 * sum;
}

KS_DEV double rc1(double e) {
    if (fabs(e) < 1e-4) return 1.0 - e * (1.0 / 3.0 - e * (0.2 - e * (1.0 / 7.0 - e * (1.0 / 9.0 - e / 11.0))));
    if (e > 0.0) { double s = sqrt(e); return atan(s) / s; }
    double s = sqrt(-e);
    return 0.5 * log((1.0 + s) / (1.0 - s)) / s;
}

KS_DEV double rj(double x, double y, double z, double p) {
    const double x0 = x, y0 = y, z0 = z;
    double A0 = (x + y + z + 2.0 * p) * 0.2;
    double delta = (p - x) * (p - y) * (p - z);
    double Q = 2000.0 * dmax(dmax(dmax(fabs(A0 - x), fabs(A0 - y)), fabs(A0 - z)), fabs(A0 - p));
    double A = A0, f = 1.0, sum = 0.0;
    for (int it = 0; it < 40 && Q * f >= fabs(A); ++it) {
        double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z), sp = sqrt(p);
        double lam = sx * sy + sy * sz + sz * sx;
        double d = (sp + sx) * (sp + sy) * (sp + sz);
        double e = (f * f * f) * delta / (d * d);
        sum += f * rc1(e) / d;
        A = (A + lam) * 0.25; x = (x + lam) * 0.25; y = (y + lam) * 0.25; z = (z + lam) * 0.25; p = (p + lam) * 0.25;
        f *= 0.25;
    }
    double Xd = f * (A0 - x0) / A, Yd = f * (A0 - y0) / A, Zd = f * (A0 - z0) / A;
    double P = -(Xd + Yd + Zd) * 0.5;
    double P2 = P * P;
    double E2 = Xd * Yd + Xd * Zd + Yd * Zd - 3.0 * P2;
    double E3 = Xd * Yd * Zd + 2.0 * E2 * P + 4.0 * P2 * P;
    double E4 = (2.0 * Xd * Yd * Zd + E2 * P + 3.0 * P2 * P) * P;
    double E5 = Xd * Yd * Zd * P2;
    double poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 - 9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return f * poly / (A * sqrt(A)) + 6.0 * sum;
}

// ---- Legendre integrals -----------------------------------------------------------------------------
KS_DEV double ellK(double m) { return rf(0.0, 1.0 - m, 1.0); }
KS_DEV double ellEc(double m) { return rf(0.0, 1.0 - m, 1.0) - m * (1.0 / 3.0) * rd(0.0, 1.0 - m, 1.0); }
KS_DEV double ellPic_small(double n, double m) { return rf(0.0, 1.0 - m, 1.0) + n * (1.0 / 3.0) * rj(0.0, 1.0 - m, 1.0, 1.0 - n); }
KS_DEV double ellPic(double n, double m) {
    if (n > 1.0) return ellK(m) - ellPic_small(m / n, m);
    return ellPic_small(n, m);
}

// F, E on |phi| <= pi/2 given (sin, cos) of the amplitude
KS_DEV double ellF_sc(double s, double c, double m) { return s * rf(c * c, 1.0 - m * s * s, 1.0); }
KS_DEV double ellE_sc(double s, double c, double m) {
    double x = c * c, y = 1.0 - m * s * s;
    return s * rf(x, y, 1.0) - m * (1.0 / 3.0) * s * s * s * rd(x, y, 1.0);
}
KS_DEV double ellPi_sc_below(double n, double s, double c, double m) {
    double s2 = s * s;
    double x = c * c, y = 1.0 - m * s2;
    return s * rf(x, y, 1.0) + n * (1.0 / 3.0) * s * s2 * rj(x, y, 1.0, 1.0 - n * s2);
}
KS_DEV double ellPi_sc(double n, double s, double c, double m) {
    double s2 = s * s;
    if (1.0 - n * s2 > 0.0) return ellPi_sc_below(n, s, c, m);
    // principal value beyond the pole, n > 1 (DLMF 19.7.9)
    double pp = sqrt((n - 1.0) * (1.0 - m / n));
    double D = sqrt(1.0 - m * s2);
    double t = s / c;
    return -ellPi_sc_below(m / n, s, c, m) + ellF_sc(s, c, m) + log(fabs((D + pp * t) / (D - pp * t))) / (2.0 * pp);
}

// unfolded to any real amplitude phi = j pi + phi'
KS_DEV double unfold(double phi, double& j) {
    j = floor(phi * (1.0 / PI) + 0.5);
    return phi - j * PI;
}
KS_DEV double ellF(double phi, double m) {
    double j; double p = unfold(phi, j);
    double s, c; sincos(p, &s, &c);
    double r = ellF_sc(s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellK(m) : r;
}
KS_DEV double ellE(double phi, double m) {
    double j; double p = unfold(phi, j);
    double s, c; sincos(p, &s, &c);
    double r = ellE_sc(s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellEc(m) : r;
}
KS_DEV double ellPi(double n, double phi, double m) {
    double j; double p = unfold(phi, j);
    double s, c; sincos(p, &s, &c);
    double r = ellPi_sc(n, s, c, m);
    return j != 0.0 ? r + 2.0 * j * ellPic(n, m) : r;
}

// ---- Jacobi ------------------------------------------------------------------------------------------
// amplitude for 0 <= m < 1 (AGM, descending Landen)
KS_DEV double am(double u, double m) {
    if (m == 0.0) return u;
    double a[10], c[10];
    a[0] = 1.0; double b = sqrt(1.0 - m); c[0] = sqrt(m);
    int N = 0;
    while (N < 9 && fabs(c[N]) > 1e-17 * a[N]) {
        double an = (a[N] + b) * 0.5;
        double bn = sqrt(a[N] * b);
        c[N + 1] = (a[N] - b) * 0.5;
        a[N + 1] = an; b = bn;
        ++N;
    }
    double phi = ldexp(a[N] * u, N);
    for (int n = N; n >= 1; --n) phi = (phi + asin(c[n] / a[n] * sin(phi))) * 0.5;
    return phi;
}

struct Jac { double sn, cn, dn; };

// (sn, cn, dn)(u|m), 0 <= m < 1
KS_DEV Jac jac_pos(double u, double m) {
    double ph = am(u, m);
    double s, c; sincos(ph, &s, &c);
    return {s, c, sqrt(1.0 - m * s * s)};
}
// any m < 1
KS_DEV Jac jac(double u, double m) {
    if (m < 0.0) {
        double s = sqrt(1.0 - m), mu = -m / (1.0 - m);
        Jac x = jac_pos(s * u, mu);
        return {x.sn / (s * x.dn), x.cn / x.dn, 1.0 / x.dn};
    }
    return jac_pos(u, m);
}
// continuous amplitude for any m < 1
KS_DEV double am_any(double u, double m) {
    if (!(m < 0.0)) return am(u, m);
    double s = sqrt(1.0 - m), mu = -m / (1.0 - m);
    double K2 = 2.0 * ellK(mu) / s;
    double j = floor(u / K2 + 0.5);
    double ur = u - j * K2;
    double ph = am(s * ur, mu);
    double sn, cn; sincos(ph, &sn, &cn);
    return atan2(sn / s, cn) + j * PI;
}

}  // namespace ksd
