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

template <class R> struct PixCT {
    R alpha, beta, lam, eta;
    int rcase, nreal;
    R k, c, fo, tau_total, tau_h;
    R r1, r3, r4, n;
    R n0, n1, d0, d1;
    R a2, b1, g0, psi0;
    R Pp, Pm, Qp, Qm, It0, Iph0;
    R up, um, f, mth, off, sgn, npol, Xth0, Gt0, Gphi0;
    bool vortical;
    R mu, scale, K2;
};
using PixC = PixCT<double>;

template <class R> struct SpinT { R a, rp, rm; };
using Spin = SpinT<double>;
template <class R>
KS_DEV SpinT<R> make_spin(R a) { R s = sqrt(1.0 - a * a); return {a, 1.0 + s, 1.0 - s}; }

// complex helpers for the quartic roots
template <class R> struct CdT { R re, im; };
using Cd = CdT<double>;
template <class R>
KS_DEV CdT<R> cadd(CdT<R> a, CdT<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
KS_DEV CdT<R> csub(CdT<R> a, CdT<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
KS_DEV CdT<R> cmul(CdT<R> a, CdT<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <class R>
KS_DEV CdT<R> cdiv(CdT<R> a, CdT<R> b) { R d = b.re * b.re + b.im * b.im; return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d}; }
template <class R>
KS_DEV R cabs_(CdT<R> a) { return sqrt(a.re * a.re + a.im * a.im); }
template <class R>
KS_DEV CdT<R> csqrt_(CdT<R> z) {
    R r = cabs_<R>(z);
    if (r == 0.0) return {0.0, 0.0};
    R s = sqrt((r + fabs(z.re)) * 0.5);
    if (z.re >= 0.0) return {s, z.im / (2.0 * s)};
    return {fabs(z.im) / (2.0 * s), copysign(s, z.im)};
}
template <class R>
KS_DEV CdT<R> ccbrt_(CdT<R> z) {
    R r = cabs_<R>(z);
    if (r == 0.0) return {0.0, 0.0};
    R th = atan2(z.im, z.re) * (1.0 / 3.0);
    R m = cbrt(r);
    R s, c; sincos(th, &s, &c);
    return {m * c, m * s};
}

// Gralla & Lupsasca (2020) eq. 95 roots of R(r) = r^4 + A r^2 + B r + C
template <class R>
KS_DEV void quartic_roots(R A, R B, R C, CdT<R> r[4]) {
    R P = -A * A / 12.0 - C;
    R Q = -A / 3.0 * ((A / 6.0) * (A / 6.0) - C) - B * B / 8.0;
    R disc = (P / 3.0) * (P / 3.0) * (P / 3.0) + (Q / 2.0) * (Q / 2.0);
    CdT<R> sd = csqrt_<R>({disc, 0.0});
    CdT<R> wp = ccbrt_<R>(cadd<R>({-Q / 2.0, 0.0}, sd));
    CdT<R> wm = cabs_<R>(wp) > 0.0 ? cdiv<R>({-P / 3.0, 0.0}, wp) : ccbrt_<R>(csub<R>({-Q / 2.0, 0.0}, sd));
    CdT<R> zz = cadd<R>(wp, wm); zz.re -= A / 3.0; zz.re *= 0.5; zz.im *= 0.5;
    CdT<R> z = csqrt_<R>(zz);
    CdT<R> Bz = cdiv<R>({B / 4.0, 0.0}, z);
    CdT<R> z2 = cmul<R>(z, z);
    CdT<R> base = {-A / 2.0 - z2.re, -z2.im};
    CdT<R> s1 = csqrt_<R>(cadd<R>(base, Bz)), s2 = csqrt_<R>(csub<R>(base, Bz));
    CdT<R> mz = {-z.re, -z.im};
    r[0] = csub<R>(mz, s1); r[1] = cadd<R>(mz, s1); r[2] = csub<R>(z, s2); r[3] = cadd<R>(z, s2);
}

template <class R>
KS_DEV void screen_to_conserved(R a, R thetao, R alpha, R beta, R& lam, R& eta) {
    R so, co; sincos(thetao, &so, &co);
    lam = -alpha * so;
    eta = (alpha * alpha - a * a) * co * co + beta * beta;
}

template <class R>
KS_DEV int count_real(const CdT<R> r[4], R& scale) {
    scale = 1.0;
    for (int i = 0; i < 4; ++i) scale = dmax(scale, cabs_<R>(r[i]));
    int n = 0;
    for (int i = 0; i < 4; ++i) n += fabs(r[i].im) <= 1e-10 * scale;
    return n;
}

template <class R>
KS_DEV int root_case_count(R a, R thetao, R alpha, R beta) {
    R lam, eta;
    screen_to_conserved<R>(a, thetao, alpha, beta, lam, eta);
    R A = a * a - eta - lam * lam, B = 2.0 * (eta + (lam - a) * (lam - a)), C = -a * a * eta;
    CdT<R> r[4];
    quartic_roots<R>(A, B, C, r);
    R s;
    return count_real<R>(r, s);
}

template <class R>
KS_DEV R r_potential(R a, R lam, R eta, R r) {
    R w = r * r + a * a - a * lam;
    return w * w - (r * r - 2.0 * r + a * a) * (eta + (lam - a) * (lam - a));
}

// ---- radial X-antiderivatives (G1 = int r dX, G2 = int r^2 dX, Gb = int dX/(r - b)) ----------------------
template <class R>
KS_DEV R V2(R nu, R s, R c, R X, R m, R Ee, R Pn) {
    R d = sqrt(1.0 - m * s * s);
    return (nu * Ee + (m - nu) * X + (2.0 * nu * m + 2.0 * nu - nu * nu - 3.0 * m) * Pn - nu * nu * s * c * d / (1.0 - nu * s * s)) /
           (2.0 * (nu - 1.0) * (m - nu));
}
template <class R>
KS_DEV R J1cn(R m0, R m1, R psi, R k) {
    R D0 = m0 * m0 - m1 * m1, Dp = D0 * k + m1 * m1;
    R nu = -m1 * m1 / D0;
    R sn = sin(psi), dn = sqrt(1.0 - k * sn * sn);
    R v = sn / dn;
    R A;
    if (D0 * Dp > 0.0) {
        A = atan(v * sqrt(Dp / D0)) / sqrt(D0 * Dp);
        if (D0 < 0.0) A = -A;
    } else {
        R q = sqrt(-D0 * Dp);
        A = -log(fabs((q + Dp * v) / (q - Dp * v))) / (2.0 * q);
    }
    return (m0 / D0) * ellPi<R>(nu, psi, k) - m1 * A;
}
template <class R>
KS_DEV R J2cn(R m0, R m1, R psi, R X, R k, R j1, R Ee) {
    R sn, cn; sincos(psi, &sn, &cn);
    R dn = sqrt(1.0 - k * sn * sn);
    R f = 1.0 / (m0 + m1 * cn);
    R m13 = m1 * m1 * m1;
    R h1 = m0 * (2.0 * k * m0 * m0 - 2.0 * k * m1 * m1 + m1 * m1) / m13;
    R h2 = -(m0 - m1) * (m0 + m1) * (k * m0 * m0 - k * m1 * m1 + m1 * m1) / m13;
    R intc2 = (Ee - (1.0 - k) * X) / k;
    return (sn * dn * f - (k / m1) * intc2 + (k * m0 * m0 / m13) * X - h1 * j1) / h2;
}
template <class R>
KS_DEV R Tn_sc(R delta, R psi, R k) {
    R j = floor(delta * (1.0 / PI) + 0.5);
    delta -= j * PI;
    R cd = cos(delta);
    R cd2 = cd * cd;
    R q = sqrt(1.0 - k * cd2);
    R sp = sin(psi);
    R w = sqrt(1.0 - k * sp * sp);
    R odd = -(log(fabs((w - q) / (w + q))) - log(fabs((1.0 - q) / (1.0 + q)))) / (2.0 * q);
    R td = tan(delta);
    R even = td != 0.0 ? -td * ellPi<R>(1.0 / cd2, psi, k) : 0.0;
    return odd + even;
}
template <class R>
KS_DEV R Sec2_sc(R delta, R psi, R X, R k, R Tn, R Ee) {
    R sn, cn; sincos(psi, &sn, &cn);
    R dn = sqrt(1.0 - k * sn * sn);
    R s2d, c2d; sincos(2.0 * delta, &s2d, &c2d);
    R int_sn2 = (X - Ee) / k;
    R int_sncn = (1.0 - dn) / k;
    R int_cos2psi = X - 2.0 * int_sn2;
    R int_sin2psi = 2.0 * int_sncn;
    R int_cos2u = int_cos2psi * c2d + int_sin2psi * s2d;
    R int_sin2u = int_sin2psi * c2d - int_cos2psi * s2d;
    R int_cosu2 = (X + int_cos2u) * 0.5;
    R cdl = cos(delta);
    R Fx = tan(psi - delta) * dn, F0 = tan(-delta);
    return (Fx - F0 - k * c2d * int_cosu2 + 0.5 * k * s2d * Tn + 0.5 * k * s2d * int_sin2u) / (1.0 - k * cdl * cdl);
}

// G2 + 2 G1 (the r^2 + 2r part of dt/dtau)
template <class R>
KS_DEV R radial_G2p2G1(const PixCT<R>& p, R X, R psi) {
    const R k = p.k;
    if (p.rcase == 2) {
        R r43 = p.r4 - p.r3;
        R s, c; sincos(psi, &s, &c);
        R Pn = ellPi<R>(p.n, psi, k);
        R Ee = ellE<R>(psi, k);
        R G1 = p.r3 * X + r43 * Pn;
        R G2 = p.r3 * p.r3 * X + 2.0 * p.r3 * r43 * Pn + r43 * r43 * V2<R>(p.n, s, c, X, k, Ee, Pn);
        return G2 + 2.0 * G1;
    } else if (p.rcase == 3) {
        R a0 = p.n1 / p.d1, a1 = (p.n0 * p.d1 - p.n1 * p.d0) / p.d1;
        R j1 = J1cn<R>(p.d0, p.d1, psi, k);
        R Ee = ellE<R>(psi, k);
        R j2 = J2cn<R>(p.d0, p.d1, psi, X, k, j1, Ee);
        R G1 = a0 * X + a1 * j1;
        R G2 = a0 * a0 * X + 2.0 * a0 * a1 * j1 + a1 * a1 * j2;
        return G2 + 2.0 * G1;
    }
    R Ee = ellE<R>(psi, k);
    R T = Tn_sc<R>(p.psi0, psi, k);
    R S2 = Sec2_sc<R>(p.psi0, psi, X, k, T, Ee);
    R G1 = p.a2 * T - p.b1 * X;
    R G2 = p.a2 * p.a2 * (S2 - X) - 2.0 * p.a2 * p.b1 * T + p.b1 * p.b1 * X;
    return G2 + 2.0 * G1;
}

template <class R>
KS_DEV R radial_Gb(const PixCT<R>& p, R X, R psi, R b) {
    const R k = p.k;
    if (p.rcase == 2) {
        R nb = p.n * (p.r3 - b) / (p.r4 - b);
        return ((p.n / nb) * X + (1.0 - p.n / nb) * ellPi<R>(nb, psi, k)) / (p.r4 - b);
    } else if (p.rcase == 3) {
        R g0 = p.n0 - b * p.d0, g1 = p.n1 - b * p.d1;
        R be0 = p.d1 / g1, be1 = (p.d0 * g1 - p.d1 * g0) / g1;
        return be0 * X + be1 * J1cn<R>(g0, g1, psi, k);
    }
    R e = p.b1 + b;
    R rho = sqrt(p.a2 * p.a2 + e * e);
    R ce = p.a2 / rho, se = e / rho;
    R ep = atan2(se, ce);
    return (-ce * Tn_sc<R>(p.psi0 + ep + 0.5 * PI, psi, k) - se * X) / rho;
}

template <class R>
KS_DEV R radial_X_of_r(const PixCT<R>& p, R r) {
    if (p.rcase == 2) {
        R sn2 = (1.0 - (p.r4 - p.r3) / (r - p.r3)) / p.n;
        return ellF<R>(asin(sqrt(sn2)), p.k);
    } else if (p.rcase == 3) {
        R cn = (p.n0 - r * p.d0) / (r * p.d1 - p.n1);
        return ellF<R>(acos(cn), p.k);
    }
    return ellF<R>(p.psi0 + atan((r + p.b1) / p.a2), p.k);
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
// PolarMarcher + QuadratureConstants, pixel_constants.jl:121-157, recurrence.jl:195-308, quadrature.jl:42-108)
template <class R>
KS_DEV PixCT<R> pixel_constants(const SpinT<R>& S, R thetao, R alpha, R beta, const double* glx, const double* glw) {
    PixCT<R> p;
    const R a = S.a;
    p.alpha = alpha; p.beta = beta;
    screen_to_conserved<R>(a, thetao, alpha, beta, p.lam, p.eta);
    R A = a * a - p.eta - p.lam * p.lam, B = 2.0 * (p.eta + (p.lam - a) * (p.lam - a)), C = -a * a * p.eta;
    CdT<R> rt[4];
    quartic_roots<R>(A, B, C, rt);
    R scale;
    int nr = count_real<R>(rt, scale);
    p.nreal = nr;
    p.r1 = p.r3 = p.r4 = p.n = p.n0 = p.n1 = p.d0 = p.d1 = p.a2 = p.b1 = p.g0 = p.psi0 = 0.0;
    if (nr == 4) {
        R v[4] = {rt[0].re, rt[1].re, rt[2].re, rt[3].re};
        // sort four values
        for (int i = 1; i < 4; ++i) for (int j = i; j > 0 && v[j] < v[j - 1]; --j) { R t = v[j]; v[j] = v[j - 1]; v[j - 1] = t; }
        R r1 = v[0], r2 = v[1], r3 = v[2], r4 = v[3];
        R r31 = r3 - r1, r32 = r3 - r2, r41 = r4 - r1, r42 = r4 - r2;
        p.k = r32 * r41 / (r31 * r42);
        p.c = sqrt(r31 * r42) * 0.5;
        p.n = r41 / r31;
        p.r1 = r1; p.r3 = r3; p.r4 = r4;
        R Xinf = ellF<R>(asin(sqrt(r31 / r41)), p.k);
        p.fo = Xinf / p.c;
        p.rcase = 2;
        if (r4 > S.rp) { p.tau_total = 2.0 * p.fo; p.tau_h = INFINITY; }
        else {
            R sh2 = r31 * (r4 - S.rp) / (r41 * (r3 - S.rp));
            p.tau_h = p.fo - ellF<R>(asin(sqrt(sh2)), p.k) / p.c;
            p.tau_total = p.tau_h;
        }
    } else if (nr == 2) {
        R re[2] = {0.0, 0.0}; CdT<R> cx = {0.0, 0.0}; int ir = 0;
        for (int i = 0; i < 4; ++i) {
            if (fabs(rt[i].im) <= 1e-10 * scale) { if (ir < 2) re[ir++] = rt[i].re; }
            else cx = rt[i];
        }
        if (re[1] < re[0]) { R t = re[0]; re[0] = re[1]; re[1] = t; }
        R r1 = re[0], r2 = re[1];
        R A_ = cabs_<R>({cx.re - r2, cx.im}), B_ = cabs_<R>({cx.re - r1, cx.im});
        R r21 = r2 - r1;
        p.k = ((A_ + B_) * (A_ + B_) - r21 * r21) / (4.0 * A_ * B_);
        p.c = sqrt(A_ * B_);
        p.n0 = B_ * r2 - A_ * r1; p.n1 = B_ * r2 + A_ * r1; p.d0 = B_ - A_; p.d1 = A_ + B_;
        p.r1 = r1;
        R Xinf = ellF<R>(acos((A_ - B_) / (A_ + B_)), p.k);
        p.fo = Xinf / p.c;
        R cnh = (p.n0 - S.rp * p.d0) / (S.rp * p.d1 - p.n1);
        p.tau_total = p.fo - ellF<R>(acos(cnh), p.k) / p.c;
        p.tau_h = INFINITY;
        p.rcase = 3;
    } else {
        CdT<R> lo = rt[0], hi = rt[3];
        for (int i = 0; i < 4; ++i) { if (rt[i].re < 0.0) lo = rt[i]; else hi = rt[i]; }
        R b2 = lo.re, a2 = fabs(lo.im), b1 = hi.re, a1 = fabs(hi.im);
        R Cc = sqrt((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2));
        R Dd = sqrt((a1 + a2) * (a1 + a2) + (b1 - b2) * (b1 - b2));
        p.k = 4.0 * Cc * Dd / ((Cc + Dd) * (Cc + Dd));
        p.c = (Cc + Dd) * 0.5;
        p.g0 = sqrt(dmax((4.0 * a2 * a2 - (Cc - Dd) * (Cc - Dd)) / ((Cc + Dd) * (Cc + Dd) - 4.0 * a2 * a2), 0.0));
        p.psi0 = atan(p.g0); p.a2 = a2; p.b1 = b1;
        R Xinf = ellF<R>(0.5 * PI + p.psi0, p.k);
        p.fo = Xinf / p.c;
        R psih = p.psi0 + atan((S.rp + b1) / a2);
        p.tau_total = p.fo - ellF<R>(psih, p.k) / p.c;
        p.tau_h = INFINITY;
        p.rcase = 4;
    }
    // ---- t~ and phi radial constants ----
    const R al = a * p.lam;
    const R dr = S.rp - S.rm;
    p.Pp = 2.0 * S.rp * (2.0 * S.rp - al) / dr;
    p.Pm = -2.0 * S.rm * (2.0 * S.rm - al) / dr;
    p.Qp = a * (2.0 * S.rp - al) / dr;
    p.Qm = -a * (2.0 * S.rm - al) / dr;
    const R rs = 4.0 * scale + 40.0;
    R Ttail;
    {
        R us = 1.0 / rs, acc = 0.0;
        for (int i = 0; i < 16; ++i) {
            R u = 0.5 * us * (1.0 + glx[i]);
            R r = 1.0 / u;
            R gt = r * r + 2.0 * r + 4.0 + p.Pp / (r - S.rp) + p.Pm / (r - S.rm);
            R h = (gt / sqrt(r_potential<R>(a, p.lam, p.eta, r)) - 1.0 - 2.0 * u) / (u * u);
            acc += glw[i] * h;
        }
        Ttail = 0.5 * us * acc - rs - 2.0 * log(rs);
    }
    const R Xinf = p.fo * p.c;
    const R Xs = radial_X_of_r<R>(p, rs);
    const R tau_s = (Xinf - Xs) / p.c;
    R psis = am<R>(Xs, p.k), psii = am<R>(Xinf, p.k);
    R Gts = radial_G2p2G1<R>(p, Xs, psis);
    R Gps = radial_Gb<R>(p, Xs, psis, S.rp), Gms = radial_Gb<R>(p, Xs, psis, S.rm);
    R Gpi = radial_Gb<R>(p, Xinf, psii, S.rp), Gmi = radial_Gb<R>(p, Xinf, psii, S.rm);
    // t~_rad(tau) = It0 - (G2 + 2G1)(X)/c + 4 tau - (Pp Gp + Pm Gm)(X)/c ; phi_rad = Iph0 - (Qp Gp + Qm Gm)(X)/c
    p.It0 = Ttail + Gts / p.c - 4.0 * tau_s + (p.Pp * Gps + p.Pm * Gms) / p.c;
    p.Iph0 = (p.Qp * Gpi + p.Qm * Gmi) / p.c;
    // ---- polar ----
    const R a2_ = a * a;
    R co = cos(thetao);
    R Dth = (1.0 - (p.eta + p.lam * p.lam) / a2_) * 0.5;
    R dsc = sqrt(Dth * Dth + p.eta / a2_);
    if (Dth < 0.0) { p.up = (p.eta / a2_) / (dsc - Dth); p.um = Dth - dsc; }
    else { p.up = Dth + dsc; p.um = -(p.eta / a2_) / (Dth + dsc); }
    p.f = 1.0 / sqrt(fabs(p.um * a2_));
    p.vortical = p.eta < 0.0;
    R sb = beta >= 0.0 ? 1.0 : -1.0;
    if (p.vortical) {
        p.mth = 1.0 - p.up / p.um;
        R sg = thetao > 0.5 * PI ? -1.0 : 1.0;
        R arg = dmin(dmax((co * co - p.um) / (p.up - p.um), 0.0), 1.0);
        R Gtho = p.f * ellF<R>(asin(sqrt(arg)), p.mth);
        p.off = sg * sb * Gtho; p.sgn = sg;
        p.npol = (p.up - p.um) / (1.0 - p.um);
    } else {
        p.mth = p.up / p.um;
        R arg = dmin(dmax(co / sqrt(p.up), -1.0), 1.0);
        R Gtho = p.f * ellF<R>(asin(arg), p.mth);
        p.off = sb * Gtho; p.sgn = sb;
        p.npol = p.up;
    }
    p.Xth0 = p.off / p.f;
    R ps0 = am_any<R>(p.Xth0, p.mth);
    p.Gt0 = p.Xth0 - ellE<R>(ps0, p.mth);
    p.Gphi0 = ellPi<R>(p.npol, ps0, p.mth);
    p.mu = -p.mth / (1.0 - p.mth);
    p.scale = sqrt(1.0 - p.mth);
    p.K2 = 2.0 * ellK<R>(p.mu) / p.scale;
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
template <class R> struct SampT { R t, r, th, ph; bool nur, nuth, ok; };
using Samp = SampT<double>;

template <class R> struct VisCT { R a, absal, al2, absbe_eps, abscos, cos2; };
using VisC = VisCT<double>;
template <class R>
KS_DEV VisCT<R> vis_constants(const PixCT<R>& p, R a, R thetao) {
    R c = cos(thetao);
    return {a, fabs(p.alpha), p.alpha * p.alpha, fabs(p.beta) + EPS, fabs(c), c * c};
}
// Krang's screen-visibility rule (recurrence.jl:273-308)
template <class R>
KS_DEV bool krang_visible(const VisCT<R>& v, R cth) {
    if (!(cth > v.abscos)) return true;
    R c2 = cth * cth, s2 = 1.0 - c2;
    R amin = v.a * sqrt(s2);
    if (!(v.absal >= amin + EPS)) return true;
    R temp = (v.cos2 - c2) * (v.al2 - v.a * v.a * s2) / (c2 - 1.0);
    R bb = sqrt(dmax(temp, 0.0));
    return !(v.absbe_eps < bb);
}

// Whether Mino time tau lies before the radial turning point fo (nu_r, direct_march.jl): tau < fo, with a sample
// within 1e-12 fo of the turning point taken as at it. With tau_total = 2 fo and N + 1 even the turning point is a
// sample in exact arithmetic, and the strict comparison would be decided by rounding (flipping the sign of the
// radial momentum there, where it is sqrt of a rounding-sized R).
KS_DEV double val(double x);     // ks_ddual.hpp (the dual scalars' val is found by argument-dependent lookup)
template <class R>
KS_DEV bool before_turning(const R& tau, const R& fo) {
    const double t = val(tau), f = val(fo);
    return t < f && (!isfinite(f) || f - t > 1e-12 * fabs(f));
}

template <class R>
KS_DEV R radius_jac(const PixCT<R>& p, R sn, R cn) {
    if (p.rcase == 2) {
        R r41 = p.r4 - p.r1, r43 = p.r4 - p.r3;
        R cn2 = r41 * cn * cn;
        return (p.r3 * cn2 - p.r1 * r43) / (cn2 - r43);
    } else if (p.rcase == 3) {
        return (p.n0 + p.n1 * cn) / (p.d0 + p.d1 * cn);
    }
    R sc = sn / cn;
    return -(p.a2 * (p.g0 - sc) / (1.0 + p.g0 * sc) + p.b1);
}

// closed-form t~ and phi at Mino time tau (the anchors)
template <class R>
KS_COLD void t_phi_direct(const PixCT<R>& p, const SpinT<R>& S, R tau, R& t, R& ph) {
    R X = p.c * (p.fo - tau);
    R psi = am<R>(X, p.k);
    R g = radial_G2p2G1<R>(p, X, psi);
    R Gp = radial_Gb<R>(p, X, psi, S.rp), Gm = radial_Gb<R>(p, X, psi, S.rm);
    R tr = p.It0 - g / p.c + 4.0 * tau - (p.Pp * Gp + p.Pm * Gm) / p.c;
    R pr = p.Iph0 - (p.Qp * Gp + p.Qm * Gm) / p.c;
    R Xt = (tau + p.off) / p.f;
    R pst = am_any<R>(Xt, p.mth);
    R Ex = Xt - ellE<R>(pst, p.mth);
    R Pi_ = ellPi<R>(p.npol, pst, p.mth);
    R Gt, Gphi;
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

template <class R>
KS_DEV void polar_from_jac(const PixCT<R>& p, R sn_t, R cn_t, R dn_t, R& th, R& cth, bool& nuth) {
    // sn_t, cn_t, dn_t at the transformed parameter mu (the recurrence's state)
    R snX = sn_t / (p.scale * dn_t);
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

template <class R>
KS_DEV SampT<R> direct_sample(const PixCT<R>& p, const SpinT<R>& S, const VisCT<R>& vis, R tau) {
    SampT<R> s;
    JacT<R> jr = jac_pos<R>(p.c * (p.fo - tau), p.k);
    s.r = radius_jac<R>(p, jr.sn, jr.cn);
    JacT<R> jt = jac_pos<R>(p.scale * (tau + p.off) / p.f, p.mu);
    R cth;
    polar_from_jac<R>(p, jt.sn, jt.cn, jt.dn, s.th, cth, s.nuth);
    t_phi_direct<R>(p, S, tau, s.t, s.ph);
    s.nur = before_turning<R>(tau, p.fo);
    s.ok = (tau <= p.tau_h) && krang_visible<R>(vis, cth);
    return s;
}

}  // namespace ksd
