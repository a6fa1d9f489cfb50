// Kerr null geodesics for the CPU oracle (test infrastructure; see ks_scalar.hpp).
//
// Restates the reference's geodesic layer (src/Geodesics/*.jl) and the closed forms it takes from
// Krang.jl (Gralla & Lupsasca 2020, "Null geodesics of the Kerr exterior"), which is not part of
// the reference tree. Units G = M = c = 1. For a pixel (alpha, beta) of an observer at
// inclination theta_o the ray is parameterized by Mino time tau from the observer (tau = 0) inward.
//
// Conventions (fixed here; Krang's own additive constants cannot be checked without Krang):
//   * lambda = -alpha sin(theta_o), eta = (alpha^2 - a^2) cos^2(theta_o) + beta^2
//     (test/highprec_reference.jl:99-100);
//   * radial potential R(r) = r^4 + A r^2 + B r + C, A = a^2 - eta - lambda^2,
//     B = 2(eta + (lambda - a)^2), C = -a^2 eta; roots by the closed form of Gralla & Lupsasca
//     (2020) eq. 95; root cases by the number of real roots, 4 / 2 / 0 (pixel_constants.jl:106-176);
//   * radial closed forms r(X), X = c (fo - tau), fo = I0_inf, per case exactly as
//     test/highprec_reference.jl:70-96 and src/Geodesics/recurrence.jl:64-110;
//   * polar closed forms (ordinary and vortical) as recurrence.jl:195-246;
//   * total Mino time: 2 I0_inf for a ray that turns outside the horizon (four real roots,
//     r4 > r+), otherwise the Mino time from the observer to the outer horizon
//     (direct_march.jl:38-45: "radial infinity or the horizon at tau_total");
//   * t~(tau) = lim_{r_o -> inf} [coordinate time elapsed from the emission point to r_o minus
//     (r_o + 2 ln r_o)], phi(tau) = integral_0^tau dphi/dtau (phi = 0 at the observer); both grow
//     with tau by the Mino-time rates, as the reference's quadrature does (quadrature.jl:276-281);
//   * flags: nu_r = tau < fo, nu_theta = sign of p^theta along the photon, ok = inside the radial
//     validity and Krang's screen-visibility rule (recurrence.jl:150-186, 273-308).
//
// The t~ and phi closed forms: with the Mino-time rates
//   dt/dtau = r^2 + 2r + 4 + P+/(r - r+) + P-/(r - r-) + a^2 cos^2(theta),
//   dphi/dtau = Q+/(r - r+) + Q-/(r - r-) + lambda / sin^2(theta),
// every radial piece is an integral of r, r^2 or 1/(r - b) over the radial Jacobi argument X, in
// which r is a Moebius function of sn^2 X (four real roots), cn X (two) or tan(am X) (none). They
// reduce to F, E and Pi (principal values where the reduction introduces a removable pole) plus
// elementary terms; the divergent r^2 + 2r part is split at a radius r_s, beyond which its
// regularized tail is a smooth integral in u = 1/r done by 16-point Gauss-Legendre. The polar
// pieces are integrals of sn^2 and 1/(1 - n sn^2) over the polar argument: E and Pi.
#pragma once
#include "ks_elliptic.hpp"

namespace kso {

// 16-point Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_16, computed once)
struct GL16 {
    double x[16], w[16];
    GL16() {
        const int n = 16;
        for (int i = 0; i < n; ++i) {
            double z = std::cos(KS_PI * (i + 0.75) / (n + 0.5)), dp = 0;
            for (int it = 0; it < 100; ++it) {
                double p1 = 1, p2 = 0;
                for (int j = 1; j <= n; ++j) { double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
                dp = n * (z * p1 - p2) / (z * z - 1);
                double dz = p1 / dp; z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            x[i] = z; w[i] = 2 / ((1 - z * z) * dp * dp);
        }
    }
};
inline const GL16& gl16() { static GL16 g; return g; }

enum : uint8_t { SAMPLE_OK = 1, SAMPLE_NUR = 2, SAMPLE_NUTH = 4 };

template <class T> struct Sample {
    T t, r, th, ph;
    bool nur, nuth, ok;
};

template <class T> struct Pixel {
    T a, alpha, beta, thetao, lam, eta;
    int nreal = 0, rcase = 0;
    Cx<T> roots[4];
    // radial
    T k, c, fo, Xinf, tau_total, tau_h;
    T r1, r2, r3, r4, n;           // four real roots
    T n0, n1, d0, d1, Aa, Bb;      // two real roots: r = (n0 + n1 cn)/(d0 + d1 cn)
    T a2, b1, g0, psi0;            // no real root: r = a2 tan(am X - psi0) - b1
    // t and phi
    T rp, rm, Pp, Pm, Qp, Qm;
    T rs, Xs, tau_s, Ttail, Gts, Gps, Gms, Gpi, Gmi;
    // polar
    T up, um, f, mth, off, sgn, npol, Xth0, Gt0, Gphi0;
    bool vortical;
};

template <class T> T r_potential(const Pixel<T>& p, T r) {
    T a = p.a;
    T w = r * r + a * a - a * p.lam;
    return w * w - (r * r - 2.0 * r + a * a) * (p.eta + (p.lam - a) * (p.lam - a));
}
template <class T> T theta_potential(T a, T eta, T lam, T th) {
    T c = kcos(th), s = ksin(th);
    return eta + a * a * c * c - lam * lam * c * c / (s * s);
}

// ---- radial X-antiderivatives ---------------------------------------------------------------------
// G1 = int_0^X r dX', G2 = int_0^X r^2 dX', Gb = int_0^X dX'/(r - b)

// V2(nu) = int_0^X dX/(1 - nu sn^2)^2 (reduction to E, F, Pi; verified against quadrature)
template <class T> T ell_V2(T nu, T psi, T X, T m) {
    T s = ksin(psi), c = kcos(psi), d = ksqrt(1.0 - m * s * s);
    return (nu * ellE(psi, m) + (m - nu) * X + (2.0 * nu * m + 2.0 * nu - nu * nu - 3.0 * m) * ellPi(nu, psi, m)
            - nu * nu * s * c * d / (1.0 - nu * s * s)) / (2.0 * (nu - 1.0) * (m - nu));
}

// J1 = int_0^X dX/(m0 + m1 cn X): even part by Pi, odd part as a function of v = sd X
template <class T> T J1cn(T m0, T m1, T psi, T k) {
    T D0 = m0 * m0 - m1 * m1, Dp = D0 * k + m1 * m1;
    T nu = -m1 * m1 / D0;
    T sn = ksin(psi), dn = ksqrt(1.0 - k * sn * sn);
    T v = sn / dn;
    T A;
    if (D0 * Dp > 0.0) {
        A = katan(v * ksqrt(Dp / D0)) / ksqrt(D0 * Dp);
        if (D0 < 0.0) A = -A;
    } else {
        T q = ksqrt(-D0 * Dp);
        A = -klog(kabs((q + Dp * v) / (q - Dp * v))) / (2.0 * q);
    }
    return (m0 / D0) * ellPi(nu, psi, k) - m1 * A;
}
// J2 = int_0^X dX/(m0 + m1 cn X)^2 by d/dX[sn dn f] = (k/m1) cn^2 - k m0^2/m1^3 + h1 f + h2 f^2
template <class T> T J2cn(T m0, T m1, T psi, T X, T k, T j1) {
    T sn = ksin(psi), cn = kcos(psi), dn = ksqrt(1.0 - k * sn * sn);
    T f = 1.0 / (m0 + m1 * cn);
    T m13 = m1 * m1 * m1;
    T h1 = m0 * (2.0 * k * m0 * m0 - 2.0 * k * m1 * m1 + m1 * m1) / m13;
    T h2 = -(m0 - m1) * (m0 + m1) * (k * m0 * m0 - k * m1 * m1 + m1 * m1) / m13;
    T intc2 = (ellE(psi, k) - (1.0 - k) * X) / k;
    return (sn * dn * f - (k / m1) * intc2 + (k * m0 * m0 / m13) * X - h1 * j1) / h2;
}
// Tn(delta) = int_0^X tan(am X' - delta) dX'
template <class T> T Tn_sc(T delta, T psi, T k) {
    double j = std::floor(val(delta) / KS_PI + 0.5);
    delta = delta - j * KS_PI;
    T cd = kcos(delta);
    T cd2 = cd * cd;
    T q = ksqrt(1.0 - k * cd2);
    T sp = ksin(psi);
    T w = ksqrt(1.0 - k * sp * sp);
    T odd = -(klog(kabs((w - q) / (w + q))) - klog(kabs((1.0 - q) / (1.0 + q)))) / (2.0 * q);
    T td = ktan(delta);
    T even = val(td) != 0.0 ? -td * ellPi(1.0 / cd2, psi, k) : T(0.0);
    return odd + even;
}
// int_0^X sec^2(am X' - delta) dX' from d/dX[tan(u) dn] (u = am X - delta), given T = Tn(delta)
template <class T> T Sec2_sc(T delta, T psi, T X, T k, T Tn) {
    T sn = ksin(psi), cn = kcos(psi), dn = ksqrt(1.0 - k * sn * sn);
    T c2d = kcos(2.0 * delta), s2d = ksin(2.0 * delta);
    T int_sn2 = (X - ellE(psi, k)) / k;
    T int_sncn = (1.0 - dn) / k;
    T int_cos2psi = X - 2.0 * int_sn2;
    T int_sin2psi = 2.0 * int_sncn;
    T int_cos2u = int_cos2psi * c2d + int_sin2psi * s2d;
    T int_sin2u = int_sin2psi * c2d - int_cos2psi * s2d;
    T int_cosu2 = (X + int_cos2u) / 2.0;
    T cdl = kcos(delta);
    T Fx = ktan(psi - delta) * dn, F0 = ktan(-delta);
    return (Fx - F0 - k * c2d * int_cosu2 + (k / 2.0) * s2d * Tn + (k / 2.0) * s2d * int_sin2u) / (1.0 - k * cdl * cdl);
}

template <class T> void radial_G12(const Pixel<T>& p, T X, T& G1, T& G2) {
    T k = p.k;
    T psi = jacobi_am(X, k);
    if (p.rcase == 2) {
        T r43 = p.r4 - p.r3;
        T Pn = ellPi(p.n, psi, k);
        G1 = p.r3 * X + r43 * Pn;
        G2 = p.r3 * p.r3 * X + 2.0 * p.r3 * r43 * Pn + r43 * r43 * ell_V2(p.n, psi, X, k);
    } else if (p.rcase == 3) {
        T a0 = p.n1 / p.d1, a1 = (p.n0 * p.d1 - p.n1 * p.d0) / p.d1;
        T j1 = J1cn(p.d0, p.d1, psi, k);
        T j2 = J2cn(p.d0, p.d1, psi, X, k, j1);
        G1 = a0 * X + a1 * j1;
        G2 = a0 * a0 * X + 2.0 * a0 * a1 * j1 + a1 * a1 * j2;
    } else {
        T Tn = Tn_sc(p.psi0, psi, k);
        T S2 = Sec2_sc(p.psi0, psi, X, k, Tn);
        G1 = p.a2 * Tn - p.b1 * X;
        G2 = p.a2 * p.a2 * (S2 - X) - 2.0 * p.a2 * p.b1 * Tn + p.b1 * p.b1 * X;
    }
}

template <class T> T radial_Gb(const Pixel<T>& p, T X, T b) {
    T k = p.k;
    T psi = jacobi_am(X, k);
    if (p.rcase == 2) {
        T nb = p.n * (p.r3 - b) / (p.r4 - b);
        return ((p.n / nb) * X + (1.0 - p.n / nb) * ellPi(nb, psi, k)) / (p.r4 - b);
    } else if (p.rcase == 3) {
        T g0 = p.n0 - b * p.d0, g1 = p.n1 - b * p.d1;
        T be0 = p.d1 / g1, be1 = (p.d0 * g1 - p.d1 * g0) / g1;
        return be0 * X + be1 * J1cn(g0, g1, psi, k);
    } else {
        T e = p.b1 + b;
        T rho = ksqrt(p.a2 * p.a2 + e * e);
        T ce = p.a2 / rho, se = e / rho;
        T ep = katan2(se, ce);
        return (-ce * Tn_sc(p.psi0 + ep + KS_PI / 2.0, psi, k) - se * X) / rho;
    }
}

// radial argument of a radius on the inbound leg
template <class T> T radial_X_of_r(const Pixel<T>& p, T r) {
    if (p.rcase == 2) {
        T sn2 = (1.0 - (p.r4 - p.r3) / (r - p.r3)) / p.n;
        return ellF(kasin(ksqrt(sn2)), p.k);
    } else if (p.rcase == 3) {
        T cn = (p.n0 - r * p.d0) / (r * p.d1 - p.n1);
        return ellF(kacos(cn), p.k);
    }
    return ellF(p.psi0 + katan((r + p.b1) / p.a2), p.k);
}

// radius from the radial argument (closed forms; recurrence.jl:160-172 for the cn^2 form)
template <class T> T radius_direct(const Pixel<T>& p, T X) {
    Jac<T> j = jacobi(X, p.k);
    if (p.rcase == 2) {
        T r41 = p.r4 - p.r1, r43 = p.r4 - p.r3;
        T cn2 = r41 * j.cn * j.cn;
        return (p.r3 * cn2 - p.r1 * r43) / (cn2 - r43);
    } else if (p.rcase == 3) {
        return (p.n0 + p.n1 * j.cn) / (p.d0 + p.d1 * j.cn);
    }
    T sc = j.sn / j.cn;
    return -(p.a2 * (p.g0 - sc) / (1.0 + p.g0 * sc) + p.b1);
}

// ---- pixel constants (K0 + K1) -------------------------------------------------------------------------
template <class T> void quartic_roots(T a, T lam, T eta, Cx<T> r[4], T& A, T& B, T& C) {
    A = a * a - eta - lam * lam;
    B = 2.0 * (eta + (lam - a) * (lam - a));
    C = -a * a * eta;
    T P = -A * A / 12.0 - C;
    T Q = -A / 3.0 * ((A / 6.0) * (A / 6.0) - C) - B * B / 8.0;
    T disc = (P / 3.0) * (P / 3.0) * (P / 3.0) + (Q / 2.0) * (Q / 2.0);
    Cx<T> sd = csqrt(Cx<T>(disc, T(0.0)));
    Cx<T> wp = ccbrt(Cx<T>(-Q / 2.0, T(0.0)) + sd);
    Cx<T> wm;
    if (val(cabs(wp)) > 0.0) wm = Cx<T>(-P / 3.0, T(0.0)) / wp;
    else wm = ccbrt(Cx<T>(-Q / 2.0, T(0.0)) - sd);
    Cx<T> z = csqrt(T(0.5) * (wp + wm - Cx<T>(A / 3.0, T(0.0))));
    Cx<T> Bz = Cx<T>(B / 4.0, T(0.0)) / z;
    Cx<T> base = Cx<T>(-A / 2.0, T(0.0)) - z * z;
    Cx<T> s1 = csqrt(base + Bz), s2 = csqrt(base - Bz);
    r[0] = -z - s1; r[1] = -z + s1; r[2] = z - s2; r[3] = z + s2;
}

// Whether a root is real: |Im r| small against the root scale (the reference counts real roots
// with Krang's `_isreal2`, a relative tolerance at the level of rounding).
template <class T> bool root_is_real(Cx<T> r, T scale) { return val(kabs(r.im)) <= 1e-10 * val(scale); }

template <class T> int num_real_roots(T a, T alpha, T beta, T thetao) {
    T lam = -alpha * ksin(thetao);
    T eta = (alpha * alpha - a * a) * kcos(thetao) * kcos(thetao) + beta * beta;
    Cx<T> r[4]; T A, B, C;
    quartic_roots(a, lam, eta, r, A, B, C);
    T scale = 1.0;
    for (int i = 0; i < 4; ++i) scale = kmax(scale, cabs(r[i]));
    int n = 0;
    for (int i = 0; i < 4; ++i) n += root_is_real(r[i], scale);
    return n;
}

template <class T> Pixel<T> pixel_constants(T a, T alpha, T beta, T thetao) {
    Pixel<T> p;
    p.a = a; p.alpha = alpha; p.beta = beta; p.thetao = thetao;
    p.lam = -alpha * ksin(thetao);
    T co = kcos(thetao);
    p.eta = (alpha * alpha - a * a) * co * co + beta * beta;
    T A, B, C;
    quartic_roots(a, p.lam, p.eta, p.roots, A, B, C);
    T scale = 1.0;
    for (int i = 0; i < 4; ++i) scale = kmax(scale, cabs(p.roots[i]));
    int nr = 0;
    for (int i = 0; i < 4; ++i) nr += root_is_real(p.roots[i], scale);
    p.nreal = nr;
    T sq = ksqrt(1.0 - a * a);
    p.rp = 1.0 + sq; p.rm = 1.0 - sq;
    // ---- radial ----
    if (nr == 4) {
        T v[4] = {p.roots[0].re, p.roots[1].re, p.roots[2].re, p.roots[3].re};
        std::sort(v, v + 4, [](T x, T y) { return x < y; });
        p.r1 = v[0]; p.r2 = v[1]; p.r3 = v[2]; p.r4 = v[3];
        T r31 = p.r3 - p.r1, r32 = p.r3 - p.r2, r41 = p.r4 - p.r1, r42 = p.r4 - p.r2;
        p.k = r32 * r41 / (r31 * r42);
        p.c = ksqrt(r31 * r42) / 2.0;
        p.n = r41 / r31;
        p.Xinf = ellF(kasin(ksqrt(r31 / r41)), p.k);
        p.fo = p.Xinf / p.c;
        p.rcase = 2;
        if (p.r4 > p.rp) {
            p.tau_total = 2.0 * p.fo; p.tau_h = T(INFINITY);
        } else {
            T sh2 = r31 * (p.r4 - p.rp) / (r41 * (p.r3 - p.rp));
            p.tau_h = p.fo - ellF(kasin(ksqrt(sh2)), p.k) / p.c;
            p.tau_total = p.tau_h;
        }
    } else if (nr == 2) {
        T re[2]; Cx<T> cx; int ir = 0;
        for (int i = 0; i < 4; ++i) {
            if (root_is_real(p.roots[i], scale)) { if (ir < 2) re[ir++] = p.roots[i].re; }
            else cx = p.roots[i];
        }
        if (re[1] < re[0]) std::swap(re[0], re[1]);
        p.r1 = re[0]; p.r2 = re[1];
        T A_ = cabs(cx - Cx<T>(p.r2)), B_ = cabs(cx - Cx<T>(p.r1));
        p.Aa = A_; p.Bb = B_;
        T r21 = p.r2 - p.r1;
        p.k = ((A_ + B_) * (A_ + B_) - r21 * r21) / (4.0 * A_ * B_);
        p.c = ksqrt(A_ * B_);
        p.n0 = B_ * p.r2 - A_ * p.r1; p.n1 = B_ * p.r2 + A_ * p.r1; p.d0 = B_ - A_; p.d1 = A_ + B_;
        p.Xinf = ellF(kacos((A_ - B_) / (A_ + B_)), p.k);
        p.fo = p.Xinf / p.c;
        T cnh = (p.n0 - p.rp * p.d0) / (p.rp * p.d1 - p.n1);
        p.tau_total = p.fo - ellF(kacos(cnh), p.k) / p.c;
        p.tau_h = T(INFINITY);
        p.rcase = 3;
    } else {
        // the pair with negative real part (-z) and the pair with positive real part (+z)
        Cx<T> lo = p.roots[0], hi = p.roots[3];
        for (int i = 0; i < 4; ++i) { if (p.roots[i].re < 0.0) lo = p.roots[i]; else hi = p.roots[i]; }
        T b2 = lo.re, a2 = kabs(lo.im), b1 = hi.re, a1 = kabs(hi.im);
        T Cc = ksqrt((a1 - a2) * (a1 - a2) + (b1 - b2) * (b1 - b2));
        T Dd = ksqrt((a1 + a2) * (a1 + a2) + (b1 - b2) * (b1 - b2));
        p.k = 4.0 * Cc * Dd / ((Cc + Dd) * (Cc + Dd));
        p.c = (Cc + Dd) / 2.0;
        p.g0 = ksqrt(kmax((4.0 * a2 * a2 - (Cc - Dd) * (Cc - Dd)) / ((Cc + Dd) * (Cc + Dd) - 4.0 * a2 * a2), T(0.0)));
        p.psi0 = katan(p.g0); p.a2 = a2; p.b1 = b1;
        p.Xinf = ellF(KS_PI / 2.0 + p.psi0, p.k);
        p.fo = p.Xinf / p.c;
        T psih = p.psi0 + katan((p.rp + b1) / a2);
        p.tau_total = p.fo - ellF(psih, p.k) / p.c;
        p.tau_h = T(INFINITY);
        p.rcase = 4;
    }
    // ---- t and phi radial constants ----
    T al = a * p.lam;
    p.Pp = 2.0 * p.rp * (2.0 * p.rp - al) / (p.rp - p.rm);
    p.Pm = -2.0 * p.rm * (2.0 * p.rm - al) / (p.rp - p.rm);
    p.Qp = a * (2.0 * p.rp - al) / (p.rp - p.rm);
    p.Qm = -a * (2.0 * p.rm - al) / (p.rp - p.rm);
    p.rs = 4.0 * scale + 40.0;
    {
        // regularized tail: int_{r_s}^inf [g_t/sqrt(R) - 1 - 2/r] dr - r_s - 2 ln r_s, in u = 1/r
        const GL16& g = gl16();
        T us = 1.0 / p.rs, acc = 0.0;
        for (int i = 0; i < 16; ++i) {
            T u = 0.5 * us * (1.0 + g.x[i]);
            T r = 1.0 / u;
            T gt = r * r + 2.0 * r + 4.0 + p.Pp / (r - p.rp) + p.Pm / (r - p.rm);
            T h = (gt / ksqrt(r_potential(p, r)) - 1.0 - 2.0 * u) / (u * u);
            acc = acc + g.w[i] * h;
        }
        p.Ttail = 0.5 * us * acc - p.rs - 2.0 * klog(p.rs);
    }
    p.Xs = radial_X_of_r(p, p.rs);
    p.tau_s = (p.Xinf - p.Xs) / p.c;
    T G1, G2;
    radial_G12(p, p.Xs, G1, G2);
    p.Gts = G2 + 2.0 * G1;
    p.Gps = radial_Gb(p, p.Xs, p.rp); p.Gms = radial_Gb(p, p.Xs, p.rm);
    p.Gpi = radial_Gb(p, p.Xinf, p.rp); p.Gmi = radial_Gb(p, p.Xinf, p.rm);
    // ---- polar ----
    T a2_ = a * a;
    T Dth = (1.0 - (p.eta + p.lam * p.lam) / a2_) / 2.0;
    T dsc = ksqrt(Dth * Dth + p.eta / a2_);
    if (Dth < 0.0) { p.up = (p.eta / a2_) / (dsc - Dth); p.um = Dth - dsc; }
    else { p.up = Dth + dsc; p.um = -(p.eta / a2_) / (Dth + dsc); }
    p.f = 1.0 / ksqrt(kabs(p.um * a2_));
    p.vortical = p.eta < 0.0;
    T sb = beta >= 0.0 ? T(1.0) : T(-1.0);
    if (p.vortical) {
        p.mth = 1.0 - p.up / p.um;
        T sg = thetao > KS_PI / 2.0 ? T(-1.0) : T(1.0);
        T arg = kmin(kmax((co * co - p.um) / (p.up - p.um), T(0.0)), T(1.0));
        T Gtho = p.f * ellF(kasin(ksqrt(arg)), p.mth);
        p.off = sg * sb * Gtho; p.sgn = sg;
        p.npol = (p.up - p.um) / (1.0 - p.um);
    } else {
        p.mth = p.up / p.um;
        T arg = kmin(kmax(co / ksqrt(p.up), T(-1.0)), T(1.0));
        T Gtho = p.f * ellF(kasin(arg), p.mth);
        p.off = sb * Gtho; p.sgn = sb;
        p.npol = p.up;
    }
    p.Xth0 = p.off / p.f;
    T psi0 = jacobi_am_any(p.Xth0, p.mth);
    p.Gt0 = p.Xth0 - ellE(psi0, p.mth);
    p.Gphi0 = ellPi(p.npol, psi0, p.mth);
    return p;
}

// ---- direct evaluation at one Mino time (the reference's `direct_sample`, Krang's closed forms) ----
template <class T> struct VisConst { T a, absal, al2, absbe_eps, abscos, cos2; };
template <class T> VisConst<T> vis_constants(const Pixel<T>& p) {
    T c = kcos(p.thetao);
    return {p.a, kabs(p.alpha), p.alpha * p.alpha, kabs(p.beta) + KS_EPS, kabs(c), c * c};
}
// Krang's screen-visibility rule (recurrence.jl:273-308)
template <class T> bool krang_visible(const VisConst<T>& v, T cth) {
    if (!(cth > v.abscos)) return true;
    T c2 = cth * cth, s2 = 1.0 - c2;
    T amin = v.a * ksqrt(s2);
    if (!(v.absal >= amin + KS_EPS)) return true;
    T temp = (v.cos2 - c2) * (v.al2 - v.a * v.a * s2) / (c2 - 1.0);
    T bb = ksqrt(kmax(temp, T(0.0)));
    return !(v.absbe_eps < bb);
}

template <class T> void polar_direct(const Pixel<T>& p, T tau, T& th, T& cth, bool& nuth) {
    T X = (tau + p.off) / p.f;
    Jac<T> j = jacobi(X, p.mth);
    if (p.vortical) {
        cth = p.sgn * ksqrt(kmax(p.um + (p.up - p.um) * j.sn * j.sn, T(0.0)));
        nuth = val(p.sgn * j.sn * j.cn) > 0.0;
    } else {
        cth = p.sgn * ksqrt(p.up) * j.sn;
        nuth = val(p.sgn * j.cn) > 0.0;
    }
    cth = kmin(kmax(cth, T(-1.0)), T(1.0));
    th = kacos(cth);
}

template <class T> void polar_integrals(const Pixel<T>& p, T tau, T& Gt, T& Gphi) {
    T X = (tau + p.off) / p.f;
    T psi = jacobi_am_any(X, p.mth);
    T Ex = X - ellE(psi, p.mth);
    if (p.vortical) {
        Gt = p.f * (p.um * (X - p.Xth0) + (p.up - p.um) * (Ex - p.Gt0) / p.mth);
        Gphi = p.f / (1.0 - p.um) * (ellPi(p.npol, psi, p.mth) - p.Gphi0);
    } else {
        Gt = p.f * p.up * (Ex - p.Gt0) / p.mth;
        Gphi = p.f * (ellPi(p.npol, psi, p.mth) - p.Gphi0);
    }
}

template <class T> void radial_t_phi(const Pixel<T>& p, T tau, T& t, T& ph) {
    T X = p.c * (p.fo - tau);
    T G1, G2;
    radial_G12(p, X, G1, G2);
    T Gp = radial_Gb(p, X, p.rp), Gm = radial_Gb(p, X, p.rm);
    t = p.Ttail + (p.Gts - (G2 + 2.0 * G1)) / p.c + 4.0 * (tau - p.tau_s) + (p.Pp * (p.Gps - Gp) + p.Pm * (p.Gms - Gm)) / p.c;
    ph = (p.Qp * (p.Gpi - Gp) + p.Qm * (p.Gmi - Gm)) / p.c;
}

// tau before the radial turning point fo (nu_r): tau < fo, with a sample within 1e-12 fo of the turning point taken
// as at it. With tau_total = 2 fo and N + 1 even the turning point is a sample in exact arithmetic, and the strict
// comparison of direct_march.jl would be decided by rounding (the radial momentum's sign, sqrt of a rounding-sized R).
template <class T> bool before_turning(const T& tau, const T& fo) {
    const double t = val(tau), f = val(fo);
    return t < f && (!std::isfinite(f) || f - t > 1e-12 * std::fabs(f));
}

template <class T> Sample<T> direct_sample(const Pixel<T>& p, T tau) {
    Sample<T> s;
    s.r = radius_direct(p, p.c * (p.fo - tau));
    T cth;
    polar_direct(p, tau, s.th, cth, s.nuth);
    T tr, pr, Gt, Gphi;
    radial_t_phi(p, tau, tr, pr);
    polar_integrals(p, tau, Gt, Gphi);
    s.t = tr + p.a * p.a * Gt;
    s.ph = pr + p.lam * Gphi;
    s.nur = before_turning(tau, p.fo);
    s.ok = (tau <= p.tau_h) && krang_visible(vis_constants(p), cth);
    return s;
}

// ---- the Recurrence(M) marcher (src/Geodesics/recurrence.jl, quadrature.jl) -----------------------------
template <class T> struct JState { T sn, cn, dn; };
template <class T> JState<T> jstate(T u, T m) { Jac<T> j = jacobi(u, m); return {j.sn, j.cn, j.dn}; }
// addition theorem (jacobi.jl:39-51)
template <class T> JState<T> jstep(const JState<T>& x, const JState<T>& d, T m) {
    T den = 1.0 - m * (x.sn * x.sn) * (d.sn * d.sn);
    return {(x.sn * d.cn * d.dn + d.sn * x.cn * x.dn) / den,
            (x.cn * d.cn - x.sn * x.dn * d.sn * d.dn) / den,
            (x.dn * d.dn - m * x.sn * x.cn * d.sn * d.cn) / den};
}
template <class T> JState<T> jstep_constants(T delta, T m) {
    JState<T> x = jstate(kabs(delta), m);
    return {kcopysign(x.sn, delta), x.cn, x.dn};
}

template <class T> struct Marcher {
    const Pixel<T>* p;
    T k;                       // radial parameter
    // polar, transformed parameter (negative_parameter_transform, jacobi.jl:69-71)
    T mu, scale, tempfac, offset, sqrt_up, um, up_m_um, sgn;
    bool vortical;
    T tau_valid;
    // quadrature constants (quadrature.jl:68-97)
    T sp, sm, Qt, Qtm, Qph, Qphm, C, nq, sqrt_eps, gc, mpp, K2;
    bool scattering;
    VisConst<T> vis;
};

template <class T> Marcher<T> make_marcher(const Pixel<T>& p) {
    Marcher<T> m;
    m.p = &p;
    m.k = p.k;
    T kth = p.mth;
    m.mu = -kth / (1.0 - kth); m.scale = ksqrt(1.0 - kth);
    m.tempfac = p.f; m.offset = p.off; m.sqrt_up = ksqrt(p.up); m.um = p.um; m.up_m_um = p.up - p.um;
    m.sgn = p.sgn; m.vortical = p.vortical;
    m.tau_valid = p.tau_h;
    T a = p.a, lam = p.lam;
    T E = p.rp * p.rp + a * a - a * lam, Em = p.rm * p.rm + a * a - a * lam;
    m.sp = -kabs(E); m.sm = -kabs(Em);
    m.Qt = (p.rp * p.rp + a * a) * E / (p.rp - p.rm);
    m.Qtm = -(p.rm * p.rm + a * a) * Em / (p.rp - p.rm);
    m.Qph = a * (2.0 * p.rp - a * lam) / (p.rp - p.rm);
    m.Qphm = -a * (2.0 * p.rm - a * lam) / (p.rp - p.rm);
    m.scattering = p.rcase == 2 && !kisfinite(p.tau_h);
    T muo = kth;                         // the (negative) parameter of the closed form
    if (p.vortical) { m.C = 1.0 / (1.0 - p.um); m.nq = (p.up - p.um) / (1.0 - p.um); }
    else { m.C = 1.0; m.nq = p.up; }
    m.K2 = 2.0 * ellK(m.mu) / m.scale;
    m.sqrt_eps = ksqrt(kmax(1.0 - m.nq, T(0.0)));
    m.gc = 1.0 / ksqrt(1.0 - muo);
    m.mpp = muo / (1.0 - muo);
    m.vis = vis_constants(p);
    return m;
}

template <class T> struct RayPoint { T r, th, cth; bool nur, nuth; T s, sn, cn, gt, gph; };

template <class T> T radius_state(const Pixel<T>& p, const JState<T>& x) {
    if (p.rcase == 2) {
        T r41 = p.r4 - p.r1, r43 = p.r4 - p.r3;
        T cn2 = r41 * x.cn * x.cn;
        return (p.r3 * cn2 - p.r1 * r43) / (cn2 - r43);
    } else if (p.rcase == 3) {
        return (p.n0 + p.n1 * x.cn) / (p.d0 + p.d1 * x.cn);
    }
    T sc = x.sn / x.cn;
    return -(p.a2 * (p.g0 - sc) / (1.0 + p.g0 * sc) + p.b1);
}

template <class T> void polar_state(const Marcher<T>& m, const JState<T>& x, T& th, bool& nuth, T& cth) {
    T snX = x.sn / (m.scale * x.dn);
    if (m.vortical) {
        cth = m.sgn * ksqrt(kmax(m.um + m.up_m_um * snX * snX, T(0.0)));
        nuth = val(m.sgn * x.sn * x.cn) > 0.0;
    } else {
        cth = m.sgn * m.sqrt_up * snX;
        nuth = val(m.sgn * x.cn) > 0.0;
    }
    cth = kmin(kmax(cth, T(-1.0)), T(1.0));
    th = kacos(cth);
}

template <class T> void smooth_integrands(const Marcher<T>& m, T r, T cth, T s, T sn, T cn, T dn, T& gt, T& gph) {
    const Pixel<T>& p = *m.p;
    T a = p.a, lam = p.lam;
    T D = r * r - 2.0 * r + a * a;
    T s2 = 1.0 - cth * cth;
    T ft = (r * r + a * a) * (r * r + a * a - a * lam) / D + a * (lam - a * s2);
    T fph = a * (2.0 * r - a * lam) / D;
    T wp = s * p.rp / (r * (r - p.rp)) / m.sp;
    T wm = s * p.rm / (r * (r - p.rm)) / m.sm;
    gt = ft - kabs(s) * (1.0 + 2.0 / r) - m.Qt * wp - m.Qtm * wm;
    T gpol = lam * m.C * (1.0 - m.gc * dn * (1.0 - m.mpp * cn * cn / 2.0)) / (1.0 - m.nq * sn * sn);
    gph = fph - m.Qph * wp - m.Qphm * wm + gpol;
}

template <class T> RayPoint<T> ray_point(const Marcher<T>& m, const JState<T>& xr, const JState<T>& xt, T tau) {
    const Pixel<T>& p = *m.p;
    RayPoint<T> q;
    q.r = radius_state(p, xr);
    polar_state(m, xt, q.th, q.nuth, q.cth);
    q.nur = before_turning(tau, p.fo);
    {   // R below 1e-14 w^2 is rounding (a sample on the radial turning point, where R = 0): s = 0 there
        const T R = r_potential(p, q.r), w = q.r * q.r + p.a * p.a - p.a * p.lam;
        q.s = (q.nur ? T(-1.0) : T(1.0)) * (R > 1e-14 * (w * w) ? ksqrt(R) : T(0.0));
    }
    q.sn = xt.sn / (m.scale * xt.dn);
    q.cn = xt.cn / xt.dn;
    T dn = 1.0 / xt.dn;
    smooth_integrands(m, q.r, q.cth, q.s, q.sn, q.cn, dn, q.gt, q.gph);
    return q;
}
template <class T> JState<T> radial_state_at(const Marcher<T>& m, T tau) { return jstate(m.p->c * (m.p->fo - tau), m.k); }
template <class T> JState<T> polar_state_at(const Marcher<T>& m, T tau) { return jstate(m.scale * (tau + m.offset) / m.tempfac, m.mu); }

template <class T> RayPoint<T> ray_point_direct(const Marcher<T>& m, T tau) {
    return ray_point(m, radial_state_at(m, tau), polar_state_at(m, tau), tau);
}
template <class T> RayPoint<T> ray_point_turning(const Marcher<T>& m) {
    const Pixel<T>& p = *m.p;
    JState<T> xt = polar_state_at(m, p.fo);
    RayPoint<T> q;
    q.r = p.r4;
    polar_state(m, xt, q.th, q.nuth, q.cth);
    q.nur = true;
    q.s = 0.0;
    q.sn = xt.sn / (m.scale * xt.dn);
    q.cn = xt.cn / xt.dn;
    T dn = 1.0 / xt.dn;
    smooth_integrands(m, q.r, q.cth, q.s, q.sn, q.cn, dn, q.gt, q.gph);
    return q;
}

template <class T> T polar_closed_form(const Marcher<T>& m, T tau, T sn, T cn) {
    T X = (tau + m.offset) / m.tempfac;
    T nw = kfloor((X + m.K2 / 2.0) / m.K2);
    T A = (katan(m.sqrt_eps * sn / cn) + nw * KS_PI) / m.sqrt_eps;
    T ph = katan(sn / cn) + nw * KS_PI;
    T eps = m.sqrt_eps * m.sqrt_eps;
    return m.p->lam * m.C * m.tempfac * m.gc * ((1.0 + m.mpp * eps / (2.0 * m.nq)) * A - m.mpp / (2.0 * m.nq) * ph);
}

template <class T> void interval_increment(const Marcher<T>& m, const RayPoint<T>& pa, const RayPoint<T>& pm, const RayPoint<T>& pb,
                                           T ta, T tb, T& It, T& Iph) {
    const Pixel<T>& p = *m.p;
    T h = tb - ta;
    T ra = pa.r, rb = pb.r;
    T dFr = (rb - ra) + 2.0 * klog(rb / ra);
    T dFp = klog((rb - p.rp) * ra / ((ra - p.rp) * rb)) / m.sp;
    T dFm = klog((rb - p.rm) * ra / ((ra - p.rm) * rb)) / m.sm;
    It = h / 6.0 * (pa.gt + 4.0 * pm.gt + pb.gt) + kabs(dFr) + m.Qt * dFp + m.Qtm * dFm;
    Iph = h / 6.0 * (pa.gph + 4.0 * pm.gph + pb.gph) + m.Qph * dFp + m.Qphm * dFm +
          polar_closed_form(m, tb, pb.sn, pb.cn) - polar_closed_form(m, ta, pa.sn, pa.cn);
}

// re-anchoring interval (recurrence.jl:359-362)
template <class T> int anchor_interval(int M, T one_minus_k) {
    double scaled = M * std::sqrt(std::max(val(one_minus_k), 0.0) / 2e-5);
    int s = (int)std::min(scaled, (double)M);
    return std::max(4, std::min(M, s));
}

inline double mino_step(double tau_total, int N) { return tau_total / (N + 1); }

// march_ray (quadrature.jl:231-296): f(k, sample, dtau) is called for k = 1..N (1-based)
template <class T, class F> void march_ray(const Pixel<T>& p, int N, int M, F&& f, T& resid_t, T& resid_ph) {
    Marcher<T> m = make_marcher(p);
    T dtau = p.tau_total / (double)(N + 1);
    T h = dtau / 2.0;
    // the split at the turning point (quadrature.jl:269, τa < fo < τb): with N + 1 even, fo = τ_total / 2 falls on a
    // sample in exact arithmetic and the strict comparison is decided by rounding; a turning point within 1e-9 dtau
    // of a sample is taken as on it (no split), the exact-arithmetic outcome
    T tol = 1e-9 * dtau;
    int Man = anchor_interval(M, 1.0 - p.k);
    JState<T> Dr = jstep_constants(-p.c * h, m.k);
    JState<T> Dt = jstep_constants(m.scale * h / m.tempfac, m.mu);
    JState<T> xr = radial_state_at(m, dtau);
    JState<T> xt = polar_state_at(m, dtau);
    RayPoint<T> pa = ray_point(m, xr, xt, dtau);
    Sample<T> d1 = direct_sample(p, dtau);
    T t = d1.t, ph = d1.ph;
    resid_t = 0.0; resid_ph = 0.0;
    f(1, Sample<T>{t, pa.r, pa.th, ph, pa.nur, pa.nuth, (dtau <= m.tau_valid) && krang_visible(m.vis, pa.cth)}, dtau);
    for (int k = 2; k <= N; ++k) {
        T ta = (double)(k - 1) * dtau, tb = (double)k * dtau;
        xr = jstep(xr, Dr, m.k);
        xt = jstep(xt, Dt, m.mu);
        RayPoint<T> pmid = ray_point(m, xr, xt, ta + h);
        bool anchor = ((k - 1) % Man) == 0;
        if (anchor) { xr = radial_state_at(m, tb); xt = polar_state_at(m, tb); }
        else { xr = jstep(xr, Dr, m.k); xt = jstep(xt, Dt, m.mu); }
        RayPoint<T> pb = ray_point(m, xr, xt, tb);
        if (m.scattering && p.fo - ta > tol && tb - p.fo > tol) {
            RayPoint<T> pt = ray_point_turning(m);
            RayPoint<T> p1 = ray_point_direct(m, (ta + p.fo) / 2.0);
            RayPoint<T> p2 = ray_point_direct(m, (p.fo + tb) / 2.0);
            T It1, Ip1, It2, Ip2;
            interval_increment(m, pa, p1, pt, ta, p.fo, It1, Ip1);
            interval_increment(m, pt, p2, pb, p.fo, tb, It2, Ip2);
            t = t + (It1 + It2);
            ph = ph + (Ip1 + Ip2);
        } else {
            T It, Ip;
            interval_increment(m, pa, pmid, pb, ta, tb, It, Ip);
            t = t + It;
            ph = ph + Ip;
        }
        if (anchor) {
            Sample<T> d = direct_sample(p, tb);
            if (d.ok) {
                resid_t = kmax(resid_t, kabs(t - d.t));
                resid_ph = kmax(resid_ph, kabs(ph - d.ph));
                t = d.t; ph = d.ph;
            }
        }
        f(k, Sample<T>{t, pb.r, pb.th, ph, pb.nur, pb.nuth, (tb <= m.tau_valid) && krang_visible(m.vis, pb.cth)}, dtau);
        pa = pb;
    }
}

// Direct marcher: closed forms at every sample (direct_march.jl:117-124)
template <class T, class F> void march_direct(const Pixel<T>& p, int N, F&& f) {
    T dtau = p.tau_total / (double)(N + 1);
    for (int k = 1; k <= N; ++k) f(k, direct_sample(p, (double)k * dtau), dtau);
}

// ---- camera and case sort -----------------------------------------------------------------------------
// Krang's screen formula lo + (hi - lo)(i - 1)/(n - 1) (camera.jl:122-124)
inline double grid_axis(double lo, double hi, int i, int n) {
    return n == 1 ? lo : lo + (hi - lo) * (double)(i) / (double)(n - 1);
}

// ---- coordinates --------------------------------------------------------------------------------------
// BL -> quasi-Cartesian Kerr-Schild (coordinates.jl:22-31)
template <class T> void quasi_cartesian_ks(T a, T r, T th, T ph, T& x, T& y, T& z) {
    T temp = ksqrt(1.0 - a * a);
    T rp = 1.0 + temp, rm = 1.0 - temp;
    T phks = ph + a / (2.0 * temp) * klog(kabs((r - rp + KS_EPS) / (r - rm + KS_EPS))) - katan2(a, r);
    T st = ksin(th), ct = kcos(th), sp = ksin(phks), cp = kcos(phks);
    x = r * st * cp; y = r * st * sp; z = r * ct;
}

}  // namespace kso
