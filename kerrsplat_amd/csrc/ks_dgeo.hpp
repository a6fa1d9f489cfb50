// Device Kerr null geodesics: per-pixel constants (K0/K1), the closed-form ("Direct") sample, and
// the Recurrence(M) marcher state (addition-theorem Jacobi steps, desingularised Simpson quadrature
// of t~ and phi, closed-form anchors every M samples).
//
// Reference: src/Geodesics/{pixel_constants,direct_march,jacobi,recurrence,quadrature}.jl, with
// the closed forms the reference takes from Krang.jl (Gralla & Lupsasca 2020) derived here:
// r is a Moebius function of sn^2 X (four real roots), cn X (two) or tan(am X) (none) of the
// radial argument X = c (fo - tau); the t~/phi anchors integrate r, r^2 and 1/(r - r+-) over X
// (Legendre F, E, Pi; principal values where a removable pole appears; elementary odd parts),
// the divergent r^2 + 2r part split at r_s with its regularised tail (u = 1/r, 16-point
// Gauss-Legendre) folded into a per-pixel constant, and the polar parts from E and Pi of the polar
// argument. Conventions: t~ = elapsed coordinate time to r_o minus r_o + 2 ln r_o (r_o -> inf),
// phi = 0 at the observer, both growing with tau by the Mino-time rates.
#pragma once
#include "ks_dmath.hpp"

namespace ksd {

// ---- per-pixel constants ------------------------------------------------------------------------------
// Field order of the SoA in HBM (one double per pixel per field, sorted slot order).
enum PixField {
    PF_ALPHA, PF_BETA, PF_LAM, PF_ETA, PF_CASE, PF_K, PF_C, PF_FO, PF_TAUTOT, PF_TAUH,
    PF_R1, PF_R3, PF_R4, PF_N,            // four real roots (also N0..D1 / A2, B1, G0, PSI0 share slots)
    PF_N0, PF_N1, PF_D0, PF_D1,
    PF_A2, PF_B1, PF_G0, PF_PSI0,
    PF_PP, PF_PM, PF_QP, PF_QM, PF_IT0, PF_IPH0,
    PF_UP, PF_UM, PF_F, PF_MTH, PF_OFF, PF_SGN, PF_NPOL, PF_XTH0, PF_GT0, PF_GPHI0, PF_VORT,
    PF_MU, PF_SCALE, PF_K2, PF_NREAL,
    PF_COUNT
};

struct PixC {
    double alpha, beta, lam, eta;
    int rcase, nreal;
    double k, c, fo, tau_total, tau_h;
    double r1, r3, r4, n;
    double n0, n1, d0, d1;
    double a2, b1, g0, psi0;
    double Pp, Pm, Qp, Qm, It0, Iph0;
    double up, um, f, mth, off, sgn, npol, Xth0, Gt0, Gphi0;
    bool vortical;
    double mu, scale, K2;
};

struct Spin { double a, rp, rm; };
KS_DEV Spin make_spin(double a) { double s = sqrt(1.0 - a * a); return {a, 1.0 + s, 1.0 - s}; }

// complex helpers for the quartic roots
struct Cd { double re, im; };
KS_DEV Cd cadd(Cd a, Cd b) { return {a.re + b.re, a.im + b.im}; }
KS_DEV Cd csub(Cd a, Cd b) { return {a.re - b.re, a.im - b.im}; }
KS_DEV Cd cmul(Cd a, Cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
KS_DEV Cd cdiv(Cd a, Cd b) { double d = b.re * b.re + b.im * b.im; return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d}; }
KS_DEV double cabs_(Cd a) { return sqrt(a.re * a.re + a.im * a.im); }
KS_DEV Cd csqrt_(Cd z) {
    double r = cabs_(z);
    if (r == 0.0) return {0.0, 0.0};
    double s = sqrt((r + fabs(z.re)) * 0.5);
    if (z.re >= 0.0) return {s, z.im / (2.0 * s)};
    return {fabs(z.im) / (2.0 * s), copysign(s, z.im)};
}
KS_DEV Cd ccbrt_(Cd z) {
    double r = cabs_(z);
    if (r == 0.0) return {0.0, 0.0};
    double th = atan2(z.im, z.re) * (1.0 / 3.0);
    double m = cbrt(r);
    double s, c; sincos(th, &s, &c);
    return {m * c, m * s};
}

// Gralla & Lupsasca (2020) eq. 95 roots of R(r) = r^4 + A r^2 + B r + C
KS_DEV void quartic_roots(double A, double B, double C, Cd r[4]) {
    double P = -A * A / 12.0 - C;
    double Q = -A / 3.0 * ((A / 6.0) * (A / 6.0) - C) - B * B / 8.0;
    double disc = (P / 3.0) * (P / 3.0) * (P / 3.0) + (Q / 2.0) * (Q / 2.0);
    Cd sd = csqrt_({disc, 0.0});
    Cd wp = ccbrt_(cadd({-Q / 2.0, 0.0}, sd));
    Cd wm = cabs_(wp) > 0.0 ? cdiv({-P / 3.0, 0.0}, wp) : ccbrt_(csub({-Q / 2.0, 0.0}, sd));
    Cd zz = cadd(wp, wm); zz.re -= A / 3.0; zz.re *= 0.5; zz.im *= 0.5;
    Cd z = csqrt_(zz);
    Cd Bz = cdiv({B / 4.0, 0.0}, z);
    Cd z2 = cmul(z, z);
    Cd base = {-A / 2.0 - z2.re, -z2.im};
    Cd s1 = csqrt_(cadd(base, Bz)), s2 = csqrt_(csub(base, Bz));
    Cd mz = {-z.re, -z.im};
    r[0] = csub(mz, s1); r[1] = cadd(mz, s1); r[2] = csub(z, s2); r[3] = cadd(z, s2);
}

KS_DEV void screen_to_conserved(double a, double thetao, double alpha, double beta, double& lam, double& eta) {
    double so, co; sincos(thetao, &so, &co);
    lam = -alpha * so;
    eta = (alpha * alpha - a * a) * co * co + beta * beta;
}

KS_DEV int count_real(const Cd r[4], double& scale) {
    scale = 1.0;
    for (int i = 0; i < 4; ++i) scale = dmax(scale, cabs_(r[i]));
    int n = 0;
    for (int i = 0; i < 4; ++i) n += fabs(r[i].im) <= 1e-10 * scale;
    return n;
}

KS_DEV int root_case_count(double a, double thetao, double alpha, double beta) {
    double lam, eta;
    screen_to_conserved(a, thetao, alpha, beta, lam, eta);
    double A = a * a - eta - lam * lam, B = 2.0 * (eta + (lam - a) * (lam - a)), C = -a * a * eta;
    Cd r[4];
    quartic_roots(A, B, C, r);
    double s;
    return count_real(r, s);
}

KS_DEV double r_potential(double a, double lam, double eta, double r) {
    double w = r * r + a * a - a * lam;
    return w * w - (r * r - 2.0 * r + a * a) * (eta + (lam - a) * (lam - a));
}

// ---- radial X-antiderivatives (G1 = int r dX, G2 = int r^2 dX, Gb = int dX/(r - b)) ----------------------
KS_DEV double V2(double nu, double s, double c, double X, double m, double Ee, double Pn) {
    double d = sqrt(1.0 - m * s * s);
    return (nu * Ee + (m - nu) * X + (2.0 * nu * m + 2.0 * nu - nu * nu - 3.0 * m) * Pn - nu * nu * s * c * d / (1.0 - nu * s * s)) /
           (2.0 * (nu - 1.0) * (m - nu));
}
KS_DEV double J1cn(double m0, double m1, double psi, double k) {
    double D0 = m0 * m0 - m1 * m1, Dp = D0 * k + m1 * m1;
    double nu = -m1 * m1 / D0;
    double sn = sin(psi), dn = sqrt(1.0 - k * sn * sn);
    double v = sn / dn;
    double A;
    if (D0 * Dp > 0.0) {
        A = atan(v * sqrt(Dp / D0)) / sqrt(D0 * Dp);
        if (D0 < 0.0) A = -A;
    } else {
        double q = sqrt(-D0 * Dp);
        A = -log(fabs((q + Dp * v) / (q - Dp * v))) / (2.0 * q);
    }
    return (m0 / D0) * ellPi(nu, psi, k) - m1 * A;
}
KS_DEV double J2cn(double m0, double m1, double psi, double X, double k, double j1, double Ee) {
    double sn, cn; sincos(psi, &sn, &cn);
    double dn = sqrt(1.0 - k * sn * sn);
    double f = 1.0 / (m0 + m1 * cn);
    double m13 = m1 * m1 * m1;
    double h1 = m0 * (2.0 * k * m0 * m0 - 2.0 * k * m1 * m1 + m1 * m1) / m13;
    double h2 = -(m0 - m1) * (m0 + m1) * (k * m0 * m0 - k * m1 * m1 + m1 * m1) / m13;
    double intc2 = (Ee - (1.0 - k) * X) / k;
    return (sn * dn * f - (k / m1) * intc2 + (k * m0 * m0 / m13) * X - h1 * j1) / h2;
}
KS_DEV double Tn_sc(double delta, double psi, double k) {
    double j = floor(delta * (1.0 / PI) + 0.5);
    delta -= j * PI;
    double cd = cos(delta);
    double cd2 = cd * cd;
    double q = sqrt(1.0 - k * cd2);
    double sp = sin(psi);
    double w = sqrt(1.0 - k * sp * sp);
    double odd = -(log(fabs((w - q) / (w + q))) - log(fabs((1.0 - q) / (1.0 + q)))) / (2.0 * q);
    double td = tan(delta);
    double even = td != 0.0 ? -td * ellPi(1.0 / cd2, psi, k) : 0.0;
    return odd + even;
}
KS_DEV double Sec2_sc(double delta, double psi, double X, double k, double Tn, double Ee) {
    double sn, cn; sincos(psi, &sn, &cn);
    double dn = sqrt(1.0 - k * sn * sn);
    double s2d, c2d; sincos(2.0 * delta, &s2d, &c2d);
    double int_sn2 = (X - Ee) / k;
    double int_sncn = (1.0 - dn) / k;
    double int_cos2psi = X - 2.0 * int_sn2;
    double int_sin2psi = 2.0 * int_sncn;
    double int_cos2u = int_cos2psi * c2d + int_sin2psi * s2d;
    double int_sin2u = int_sin2psi * c2d - int_cos2psi * s2d;
    double int_cosu2 = (X + int_cos2u) * 0.5;
    double cdl = cos(delta);
    double Fx = tan(psi - delta) * dn, F0 = tan(-delta);
    return (Fx - F0 - k * c2d * int_cosu2 + 0.5 * k * s2d * Tn + 0.5 * k * s2d * int_sin2u) / (1.0 - k * cdl * cdl);
}

// G2 + 2 G1 (the r^2 + 2r part of dt/dtau)
KS_DEV double radial_G2p2G1(const PixC& p, double X, double psi) {
    const double k = p.k;
    if (p.rcase == 2) {
        double r43 = p.r4 - p.r3;
        double s, c; sincos(psi, &s, &c);
        double Pn = ellPi(p.n, psi, k);
        double Ee = ellE(psi, k);
        double G1 = p.r3 * X + r43 * Pn;
        double G2 = p.r3 * p.r3 * X + 2.0 * p.r3 * r43 * Pn + r43 * r43 * V2(p.n, s, c, X, k, Ee, Pn);
        return G2 + 2.0 * G1;
    } else if (p.rcase == 3) {
        double a0 = p.n1 / p.d1, a1 = (p.n0 * p.d1 - p.n1 * p.d0) / p.d1;
        double j1 = J1cn(p.d0, p.d1, psi, k);
        double Ee = ellE(psi, k);
        double j2 = J2cn(p.d0, p.d1, psi, X, k, j1, Ee);
        double G1 = a0 * X + a1 * j1;
        double G2 = a0 * a0 * X + 2.0 * a0 * a1 * j1 + a1 * a1 * j2;
        return G2 + 2.0 * G1;
    }
    double Ee = ellE(psi, k);
    double T = Tn_sc(p.psi0, psi, k);
    double S2 = Sec2_sc(p.psi0, psi, X, k, T, Ee);
    double G1 = p.a2 * T - p.b1 * X;
    double G2 = p.a2 * p.a2 * (S2 - X) - 2.0 * p.a2 * p.b1 * T + p.b1 * p.b1 * X;
    return G2 + 2.0 * G1;
}

KS_DEV double radial_Gb(const PixC& p, double X, double psi, double b) {
    const double k = p.k;
    if (p.rcase == 2) {
        double nb = p.n * (p.r3 - b) / (p.r4 - b);
        return ((p.n / nb) * X + (1.0 - p.n / nb) * ellPi(nb, psi, k)) / (p.r4 - b);
    } else if (p.rcase == 3) {
        double g0 = p.n0 - b * p.d0, g1 = p.n1 - b * p.d1;
        double be0 = p.d1 / g1, be1 = (p.d0 * g1 - p.d1 * g0) / g1;
        return be0 * X + be1 * J1cn(g0, g1, psi, k);
    }
    double e = p.b1 + b;
    double rho = sqrt(p.a2 * p.a2 + e * e);
    double ce = p.a2 / rho, se = e / rho;
    double ep = atan2(se, ce);
    return (-ce * Tn_sc(p.psi0 + ep + 0.5 * PI, psi, k) - se * X) / rho;
}

KS_DEV double radial_X_of_r(const PixC& p, double r) {
    if (p.rcase == 2) {
        double sn2 = (1.0 - (p.r4 - p.r3) / (r - p.r3)) / p.n;
        return ellF(asin(sqrt(sn2)), p.k);
    } else if (p.rcase == 3) {
        double cn = (p.n0 - r * p.d0) / (r * p.d1 - p.n1);
        return ellF(acos(cn), p.k);
    }
    return ellF(p.psi0 + atan((r + p.b1) / p.a2), p.k);
}

// 16-point Gauss-Legendre nodes and weights on [-1, 1]
__constant__ double GL16_X[16] = {-0.9894009349916499, -0.9445750230732326, -0.8656312023878318, -0.7554044083550030,
                                  -0.6178762444026438, -0.4580167776572274, -0.2816035507792589, -0.0950125098376374,
                                   0.0950125098376374,  0.2816035507792589,  0.4580167776572274,  0.6178762444026438,
                                   0.7554044083550030,  0.8656312023878318,  0.9445750230732326,  0.9894009349916499};
__constant__ double GL16_W[16] = {0.0271524594117541, 0.0622535239386479, 0.0951585116824928, 0.1246289712555339,
                                  0.1495959888165767, 0.1691565193950025, 0.1826034150449236, 0.1894506104550685,
                                  0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                                  0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};

// K1: the constants of one pixel (the reference's Krang.SlowLightIntensityPixel + radial_marcher +
// PolarMarcher + QuadratureConstants, pixel_constants.jl:121-157, recurrence.jl:273-446)
KS_DEV PixC pixel_constants(const Spin& S, double thetao, double alpha, double beta, const double* glx, const double* glw) {
    PixC p;
    const double a = S.a;
    p.alpha = alpha; p.beta = beta;
    screen_to_conserved(a, thetao, alpha, beta, p.lam, p.eta);
    double A = a * a - p.eta - p.lam * p.lam, B = 2.0 * (p.eta + (p.lam - a) * (p.lam - a)), C = -a * a * p.eta;
    Cd rt[4];
    quartic_roots(A, B, C, rt);
    double scale;
    int nr = count_real(rt, scale);
    p.nreal = nr;
    p.r1 = p.r3 = p.r4 = p.n = p.n0 = p.n1 = p.d0 = p.d1 = p.a2 = p.b1 = p.g0 = p.psi0 = 0.0;
    if (nr == 4) {
        double v[4] = {rt[0].re, rt[1].re, rt[2].re, rt[3].re};
        // sort four values
        for (int i = 1; i < 4; ++i) for (int j = i; j > 0 && v[j] < v[j - 1]; --j) { double t = v[j]; v[j] = v[j - 1]; v[j - 1] = t; }
        double r1 = v[0], r2 = v[1], r3 = v[2], r4 = v[3];
        double r31 = r3 - r1, r32 = r3 - r2, r41 = r4 - r1, r42 = r4 - r2;
        p.k = r32 * r41 / (r31 * r42);
        p.c = sqrt(r31 * r42) * 0.5;
        p.n = r41 / r31;
        p.r1 = r1; p.r3 = r3; p.r4 = r4;
        double Xinf = ellF(asin(sqrt(r31 / r41)), p.k);
        p.fo = Xinf / p.c;
        p.rcase = 2;
        if (r4 > S.rp) { p.tau_total = 2.0 * p.fo; p.tau_h = INFINITY; }
        else {
            double sh2 = r31 * (r4 - S.rp) / (r41 * (r3 - S.rp));
            p.tau_h = p.fo - ellF(asin(sqrt(sh2)), p.k) / p.c;
            p.tau_total = p.tau_h;
        }
    } else if (nr == 2) {
        double re[2] = {0.0, 0.0}; Cd cx = {0.0, 0.0}; int ir = 0;
        for (int i = 0; i < 4; ++i) {
            if (fabs(rt[i].im) <= 1e-10 * scale) { if (ir < 2) re[ir++] = rt[i].re; }
            else cx = rt[i];
        }
        if (re[1] < re[0]) { double t = re[0]; re[0] = re[1]; re[1] = t; }
        double r1 = re[0], r2 = re[1];
        double A_ = cabs_({cx.re - r2, cx.im}), B_ = cabs_({cx.re - r1, cx.im});
        double r21 = r2 - r1;
        p.k = ((A_ + B_) * (A_ + B_) - r21 * r21) / (4.0 * A_ * B_);
        p.c = sqrt(A_ * B_);
        p.n0 = B_ * r2 - A_ * r1; p.n1 = B_ * r2 + A_ * r1; p.d0 = B_ - A_; p.d1 = A_ + B_;
        p.r1 = r1;
        double Xinf = ellF(acos((A_ - B_) / (A_ + B_)), p.k);
        p.fo = Xinf / p.c;
        double cnh = (p.n0 - S.rp * p.d0) / (S.rp * p.d1 - p.n1);
        p.tau_total = p.fo - ellF(acos(cnh), p.k) / p.c;
        p.tau_h = INFINITY;
        p.rcase = 3;
    } else {
        Cd lo = rt[0], hi = rt[3];
        for (int i = 0; i < 4; ++i) { if (rt[i].re < 0.0) lo = rt[i]; else hi = rt[i]; }
        double b2 = lo.re, a2 = fabs(lo.im), b1 = hi.re, a1 = fabs(hi.im);
        double Cc = sqrt((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2));
        double Dd = sqrt((a1 + a2) * (a1 + a2) + (b1 - b2) * (b1 - b2));
        p.k = 4.0 * Cc * Dd / ((Cc + Dd) * (Cc + Dd));
        p.c = (Cc + Dd) * 0.5;
        p.g0 = sqrt(dmax((4.0 * a2 * a2 - (Cc - Dd) * (Cc - Dd)) / ((Cc + Dd) * (Cc + Dd) - 4.0 * a2 * a2), 0.0));
        p.psi0 = atan(p.g0); p.a2 = a2; p.b1 = b1;
        double Xinf = ellF(0.5 * PI + p.psi0, p.k);
        p.fo = Xinf / p.c;
        double psih = p.psi0 + atan((S.rp + b1) / a2);
        p.tau_total = p.fo - ellF(psih, p.k) / p.c;
        p.tau_h = INFINITY;
        p.rcase = 4;
    }
    // ---- t~ and phi radial constants ----
    const double al = a * p.lam;
    const double dr = S.rp - S.rm;
    p.Pp = 2.0 * S.rp * (2.0 * S.rp - al) / dr;
    p.Pm = -2.0 * S.rm * (2.0 * S.rm - al) / dr;
    p.Qp = a * (2.0 * S.rp - al) / dr;
    p.Qm = -a * (2.0 * S.rm - al) / dr;
    const double rs = 4.0 * scale + 40.0;
    double Ttail;
    {
        double us = 1.0 / rs, acc = 0.0;
        for (int i = 0; i < 16; ++i) {
            double u = 0.5 * us * (1.0 + glx[i]);
            double r = 1.0 / u;
            double gt = r * r + 2.0 * r + 4.0 + p.Pp / (r - S.rp) + p.Pm / (r - S.rm);
            double h = (gt / sqrt(r_potential(a, p.lam, p.eta, r)) - 1.0 - 2.0 * u) / (u * u);
            acc += glw[i] * h;
        }
        Ttail = 0.5 * us * acc - rs - 2.0 * log(rs);
    }
    const double Xinf = p.fo * p.c;
    const double Xs = radial_X_of_r(p, rs);
    const double tau_s = (Xinf - Xs) / p.c;
    double psis = am(Xs, p.k), psii = am(Xinf, p.k);
    double Gts = radial_G2p2G1(p, Xs, psis);
    double Gps = radial_Gb(p, Xs, psis, S.rp), Gms = radial_Gb(p, Xs, psis, S.rm);
    double Gpi = radial_Gb(p, Xinf, psii, S.rp), Gmi = radial_Gb(p, Xinf, psii, S.rm);
    // t~_rad(tau) = It0 - (G2 + 2G1)(X)/c + 4 tau - (Pp Gp + Pm Gm)(X)/c ; phi_rad = Iph0 - (Qp Gp + Qm Gm)(X)/c
    p.It0 = Ttail + Gts / p.c - 4.0 * tau_s + (p.Pp * Gps + p.Pm * Gms) / p.c;
    p.Iph0 = (p.Qp * Gpi + p.Qm * Gmi) / p.c;
    // ---- polar ----
    const double a2_ = a * a;
    double co = cos(thetao);
    double Dth = (1.0 - (p.eta + p.lam * p.lam) / a2_) * 0.5;
    double dsc = sqrt(Dth * Dth + p.eta / a2_);
    if (Dth < 0.0) { p.up = (p.eta / a2_) / (dsc - Dth); p.um = Dth - dsc; }
    else { p.up = Dth + dsc; p.um = -(p.eta / a2_) / (Dth + dsc); }
    p.f = 1.0 / sqrt(fabs(p.um * a2_));
    p.vortical = p.eta < 0.0;
    double sb = beta >= 0.0 ? 1.0 : -1.0;
    if (p.vortical) {
        p.mth = 1.0 - p.up / p.um;
        double sg = thetao > 0.5 * PI ? -1.0 : 1.0;
        double arg = dmin(dmax((co * co - p.um) / (p.up - p.um), 0.0), 1.0);
        double Gtho = p.f * ellF(asin(sqrt(arg)), p.mth);
        p.off = sg * sb * Gtho; p.sgn = sg;
        p.npol = (p.up - p.um) / (1.0 - p.um);
    } else {
        p.mth = p.up / p.um;
        double arg = dmin(dmax(co / sqrt(p.up), -1.0), 1.0);
        double Gtho = p.f * ellF(asin(arg), p.mth);
        p.off = sb * Gtho; p.sgn = sb;
        p.npol = p.up;
    }
    p.Xth0 = p.off / p.f;
    double ps0 = am_any(p.Xth0, p.mth);
    p.Gt0 = p.Xth0 - ellE(ps0, p.mth);
    p.Gphi0 = ellPi(p.npol, ps0, p.mth);
    p.mu = -p.mth / (1.0 - p.mth);
    p.scale = sqrt(1.0 - p.mth);
    p.K2 = 2.0 * ellK(p.mu) / p.scale;
    return p;
}

KS_DEV void store_pixel(double* soa, int64_t stride, int64_t j, const PixC& p) {
    double* q = soa + j;
#define ST(F, V) q[(int64_t)(F) * stride] = (V)
    ST(PF_ALPHA, p.alpha); ST(PF_BETA, p.beta); ST(PF_LAM, p.lam); ST(PF_ETA, p.eta); ST(PF_CASE, (double)p.rcase);
    ST(PF_K, p.k); ST(PF_C, p.c); ST(PF_FO, p.fo); ST(PF_TAUTOT, p.tau_total); ST(PF_TAUH, p.tau_h);
    ST(PF_R1, p.r1); ST(PF_R3, p.r3); ST(PF_R4, p.r4); ST(PF_N, p.n);
    ST(PF_N0, p.n0); ST(PF_N1, p.n1); ST(PF_D0, p.d0); ST(PF_D1, p.d1);
    ST(PF_A2, p.a2); ST(PF_B1, p.b1); ST(PF_G0, p.g0); ST(PF_PSI0, p.psi0);
    ST(PF_PP, p.Pp); ST(PF_PM, p.Pm); ST(PF_QP, p.Qp); ST(PF_QM, p.Qm); ST(PF_IT0, p.It0); ST(PF_IPH0, p.Iph0);
    ST(PF_UP, p.up); ST(PF_UM, p.um); ST(PF_F, p.f); ST(PF_MTH, p.mth); ST(PF_OFF, p.off); ST(PF_SGN, p.sgn);
    ST(PF_NPOL, p.npol); ST(PF_XTH0, p.Xth0); ST(PF_GT0, p.Gt0); ST(PF_GPHI0, p.Gphi0); ST(PF_VORT, p.vortical ? 1.0 : 0.0);
    ST(PF_MU, p.mu); ST(PF_SCALE, p.scale); ST(PF_K2, p.K2); ST(PF_NREAL, (double)p.nreal);
#undef ST
}
KS_DEV PixC load_pixel(const double* __restrict__ soa, int64_t stride, int64_t j) {
    PixC p;
    const double* q = soa + j;
#define LD(F) q[(int64_t)(F) * stride]
    p.alpha = LD(PF_ALPHA); p.beta = LD(PF_BETA); p.lam = LD(PF_LAM); p.eta = LD(PF_ETA); p.rcase = (int)LD(PF_CASE);
    p.k = LD(PF_K); p.c = LD(PF_C); p.fo = LD(PF_FO); p.tau_total = LD(PF_TAUTOT); p.tau_h = LD(PF_TAUH);
    p.r1 = LD(PF_R1); p.r3 = LD(PF_R3); p.r4 = LD(PF_R4); p.n = LD(PF_N);
    p.n0 = LD(PF_N0); p.n1 = LD(PF_N1); p.d0 = LD(PF_D0); p.d1 = LD(PF_D1);
    p.a2 = LD(PF_A2); p.b1 = LD(PF_B1); p.g0 = LD(PF_G0); p.psi0 = LD(PF_PSI0);
    p.Pp = LD(PF_PP); p.Pm = LD(PF_PM); p.Qp = LD(PF_QP); p.Qm = LD(PF_QM); p.It0 = LD(PF_IT0); p.Iph0 = LD(PF_IPH0);
    p.up = LD(PF_UP); p.um = LD(PF_UM); p.f = LD(PF_F); p.mth = LD(PF_MTH); p.off = LD(PF_OFF); p.sgn = LD(PF_SGN);
    p.npol = LD(PF_NPOL); p.Xth0 = LD(PF_XTH0); p.Gt0 = LD(PF_GT0); p.Gphi0 = LD(PF_GPHI0); p.vortical = LD(PF_VORT) != 0.0;
    p.mu = LD(PF_MU); p.scale = LD(PF_SCALE); p.K2 = LD(PF_K2); p.nreal = (int)LD(PF_NREAL);
#undef LD
    return p;
}

// ---- sample-level closed forms ------------------------------------------------------------------------------
struct Samp { double t, r, th, ph; bool nur, nuth, ok; };

struct VisC { double a, absal, al2, absbe_eps, abscos, cos2; };
KS_DEV VisC vis_constants(const PixC& p, double a, double thetao) {
    double c = cos(thetao);
    return {a, fabs(p.alpha), p.alpha * p.alpha, fabs(p.beta) + EPS, fabs(c), c * c};
}
// Krang's screen-visibility rule (recurrence.jl:508-517)
KS_DEV bool krang_visible(const VisC& v, double cth) {
    if (!(cth > v.abscos)) return true;
    double c2 = cth * cth, s2 = 1.0 - c2;
    double amin = v.a * sqrt(s2);
    if (!(v.absal >= amin + EPS)) return true;
    double temp = (v.cos2 - c2) * (v.al2 - v.a * v.a * s2) / (c2 - 1.0);
    double bb = sqrt(dmax(temp, 0.0));
    return !(v.absbe_eps < bb);
}

KS_DEV double radius_jac(const PixC& p, double sn, double cn) {
    if (p.rcase == 2) {
        double r41 = p.r4 - p.r1, r43 = p.r4 - p.r3;
        double cn2 = r41 * cn * cn;
        return (p.r3 * cn2 - p.r1 * r43) / (cn2 - r43);
    } else if (p.rcase == 3) {
        return (p.n0 + p.n1 * cn) / (p.d0 + p.d1 * cn);
    }
    double sc = sn / cn;
    return -(p.a2 * (p.g0 - sc) / (1.0 + p.g0 * sc) + p.b1);
}

// closed-form t~ and phi at Mino time tau (the anchors)
KS_COLD void t_phi_direct(const PixC& p, const Spin& S, double tau, double& t, double& ph) {
    double X = p.c * (p.fo - tau);
    double psi = am(X, p.k);
    double g = radial_G2p2G1(p, X, psi);
    double Gp = radial_Gb(p, X, psi, S.rp), Gm = radial_Gb(p, X, psi, S.rm);
    double tr = p.It0 - g / p.c + 4.0 * tau - (p.Pp * Gp + p.Pm * Gm) / p.c;
    double pr = p.Iph0 - (p.Qp * Gp + p.Qm * Gm) / p.c;
    double Xt = (tau + p.off) / p.f;
    double pst = am_any(Xt, p.mth);
    double Ex = Xt - ellE(pst, p.mth);
    double Pi_ = ellPi(p.npol, pst, p.mth);
    double Gt, Gphi;
    if (p.vortical) {
        Gt = p.f * (p.um * (Xt - p.Xth0) + (p.up - p.um) * (Ex - p.Gt0) / p.mth);
        Gphi = p.f / (1.0 - p.um) * (Pi_ - p.Gphi0);
    } else {
        Gt = p.f * p.up * (Ex - p.Gt0) / p.mth;
        Gphi = p.f * (Pi_ - p.Gphi0);
    }
    t = tr + S.a * S.a * Gt;
    ph = pr + p.lam * Gphi;
}

KS_DEV void polar_from_jac(const PixC& p, double sn_t, double cn_t, double dn_t, double& th, double& cth, bool& nuth) {
    // sn_t, cn_t, dn_t at the transformed parameter mu (the recurrence's state)
    double snX = sn_t / (p.scale * dn_t);
    if (p.vortical) {
        cth = p.sgn * sqrt(dmax(p.um + (p.up - p.um) * snX * snX, 0.0));
        nuth = p.sgn * sn_t * cn_t > 0.0;
    } else {
        cth = p.sgn * sqrt(p.up) * snX;
        nuth = p.sgn * cn_t > 0.0;
    }
    cth = dmin(dmax(cth, -1.0), 1.0);
    th = acos(cth);
}

KS_DEV Samp direct_sample(const PixC& p, const Spin& S, const VisC& vis, double tau) {
    Samp s;
    Jac jr = jac_pos(p.c * (p.fo - tau), p.k);
    s.r = radius_jac(p, jr.sn, jr.cn);
    Jac jt = jac_pos(p.scale * (tau + p.off) / p.f, p.mu);
    double cth;
    polar_from_jac(p, jt.sn, jt.cn, jt.dn, s.th, cth, s.nuth);
    t_phi_direct(p, S, tau, s.t, s.ph);
    s.nur = tau < p.fo;
    s.ok = (tau <= p.tau_h) && krang_visible(vis, cth);
    return s;
}

}  // namespace ksd
