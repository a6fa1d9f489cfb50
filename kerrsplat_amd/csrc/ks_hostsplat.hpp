// Host-side per-splat constants of a model (the work of ks_model_create before the upload): modified Bessel
// functions, the kappa 2F1 factor and digamma for the population tables and their partials, and the constants of
// one splat in FP64 with their copies in another scalar. Shared by the library and the host-side work model.
// Reference: polarized.jl:41-108 (the splat's rows), populations.jl:13-104, kappa.jl:13-16, trajectories.jl:80-128.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include "ks_dadjoint.hpp"

namespace ksd {

inline double host_besselk(double x, double& k0, double& k1) {
    // K_nu(x) = int_0^inf exp(-x cosh s) cosh(nu s) ds, trapezoidal rule (exponentially convergent)
    const double h = 0.125;
    double s0 = 0.5 * std::exp(-x), s1 = s0;
    for (int i = 1; i < 4000; ++i) {
        double s = h * i;
        double e = std::exp(-x * std::cosh(s));
        s0 += e; s1 += e * std::cosh(s);
        if (e < 1e-40 * s0) break;
    }
    k0 = h * s0; k1 = h * s1;
    return k0 + 2.0 * k1 / x;
}
// (the scalar a double, or duals whose partials the series carries term by term: the series stops when the
// term is negligible in the value and in every partial)
inline bool series_done(double t, double s) { return std::fabs(t) < 1e-17 * std::fabs(s); }
template <int N> inline bool series_done(const DN<N, double>& t, const DN<N, double>& s) {
    bool done = series_done(t.v, s.v);
    for (int i = 0; i < N; ++i) done = done && std::fabs(t.d[i]) <= 1e-17 * std::fabs(s.d[i]);
    return done;
}
template <class T> inline T host_hyp2f1_neg(T a, T b, T c, T z) {
    T w = z / (z - 1.0), bb = c - b, term(1.0), sum(1.0);
    for (int n = 0; n < 200000; ++n) {
        term = term * ((a + (double)n) * (bb + (double)n) / ((c + (double)n) * (n + 1.0)) * w);
        sum = sum + term;
        if (series_done(term, sum)) break;
    }
    return pow(1.0 - z, -a) * sum;
}

// digamma for x > 0: the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 10, then the asymptotic series
// ln x - 1/(2x) - sum B_2k / (2k x^2k) (DLMF 5.5.2, 5.11.2)
inline double host_digamma(double x) {
    double r = 0.0;
    while (x < 10.0) { r -= 1.0 / x; x += 1.0; }
    const double i2 = 1.0 / (x * x);
    const double ser = i2 * (1.0 / 12.0 - i2 * (1.0 / 120.0 - i2 * (1.0 / 252.0 - i2 * (1.0 / 240.0 - i2 * (1.0 / 132.0 - i2 * (691.0 / 32760.0 - i2 / 12.0))))));
    return r + std::log(x) - 0.5 / x - ser;
}
// kappa's 2F1 factor 2F1(k - 1/3, k + 1; k + 2/3; -k w) (kappa.jl:13-16)
template <class T> inline T host_kappa_hyp(T k, T w) { return host_hyp2f1_neg(k - 1.0 / 3.0, k + 1.0, k + 2.0 / 3.0, -(k * w)); }
// d/d ln Theta (or d/d ln w) of K0/K2 and K1/K2 at x = 1/Theta: -x d/dx with K0' = -K1, K1' = -K0 - K1/x,
// K2 = K0 + 2 K1/x
inline void host_bessel_ratio_dlog(double xb, double k0, double k1, double k2, double& dk0k2, double& dk1k2) {
    const double dk1 = -k0 - k1 / xb, dk2 = -k1 - 2.0 * k0 / xb - 4.0 * k1 / (xb * xb);
    dk0k2 = -xb * (-k1 * k2 - k0 * dk2) / (k2 * k2);
    dk1k2 = -xb * (dk1 * k2 - k1 * dk2) / (k2 * k2);
}

// Copies of the per-splat constants in another scalar (FP32 for the mixed path; the work model's counting
// scalar), every field converted
template <class T> inline void splat_as(const SplatC& a, SplatCT<T>& b) {
    std::memset((void*)&b, 0, sizeof b);
    b.x = T(a.x); b.y = T(a.y); b.z = T(a.z);
    for (int i = 0; i < 9; ++i) b.R[i] = T(a.R[i]);
    for (int i = 0; i < 3; ++i) { b.is[i] = T(a.is[i]); b.bh[i] = T(a.bh[i]); b.u[i] = T(a.u[i]); }
    b.t0 = T(a.t0); b.iw = T(a.iw); b.omega = T(a.omega); b.rsup2 = T(a.rsup2); b.lnA = T(a.lnA); b.Bmag = T(a.Bmag);
    b.p1 = T(a.p1); b.p2 = T(a.p2); b.p3 = T(a.p3);
    for (int i = 0; i < 6; ++i) b.g[i] = T(a.g[i]);
    for (int i = 0; i < 3; ++i) b.kc[i] = T(a.kc[i]);
    b.kr = T(a.kr);
    for (int f = 0; f < KK_N; ++f) b.kk.v[f] = T(a.kk.v[f]);
    b.knot0 = a.knot0; b.nk = a.nk; b.kt0 = a.kt0; b.kind = a.kind;
}
template <class T> inline void adj_as(const SplatAdj& a, SplatAdjT<T>& b) {
    std::memset((void*)&b, 0, sizeof b);
    for (int i = 0; i < 4; ++i) b.qn[i] = T(a.qn[i]);
    b.iqn = T(a.iqn); b.dk0k2 = T(a.dk0k2); b.dk1k2 = T(a.dk1k2);
    b.sthB = T(a.sthB); b.cthB = T(a.cthB); b.sphB = T(a.sphB); b.cphB = T(a.cphB);
    for (int i = 0; i < 8; ++i) b.dpop[i] = T(a.dpop[i]);
    for (int q = 0; q < 3; ++q)
        for (int f = 0; f < KK_N; ++f) b.dkk[q][f] = T(a.dkk[q][f]);
}

// The host part of ks_model_create for splat i (its parameter column p, nrows rows): the constants the kernels
// read (SplatC: rotation, inverse scales, envelope, support radius, field, velocity, the population's Bessel, Gamma
// and 2F1 factors, kappa's fit factors, the knot sphere) and those of the adjoint (SplatAdj, with the partials of
// the per-splat tables in the population rows). kind is the model's population (ModelKind).
inline void host_splat_constants(int kind, int nrows, int i, const double* p, double gamma_max, bool with_knots,
                                 const double* knots, int nknots, SplatC& c, SplatAdj& sa) {
    const double SR2 = -2.0 * std::log(1e-6);
    std::memset((void*)&sa, 0, sizeof sa);
    std::memset((void*)&c, 0, sizeof c);
    c.x = p[0]; c.y = p[1]; c.z = p[2];
    double q1 = p[6], q2 = p[7], q3 = p[8], q4 = p[9];
    double nq = std::sqrt(q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4);
    double w = q1 / nq, x = q2 / nq, y = q3 / nq, z = q4 / nq;
    double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                   2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                   2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
    std::memcpy(c.R, R, sizeof R);
    for (int a_ = 0; a_ < 3; ++a_) c.is[a_] = std::exp(-p[3 + a_]);
    c.t0 = p[10]; c.iw = std::exp(-p[11]);
    c.omega = p[nrows - 1];
    double smax = std::max(std::max(p[3], p[4]), p[5]);
    c.rsup2 = SR2 * std::exp(2.0 * smax);
    c.lnA = p[12];
    c.kind = kind;
    if (with_knots) {
        c.knot0 = nknots * i; c.nk = nknots; c.kt0 = 0;
        // the sphere holding the knots (the centre stays in their convex hull)
        const double* K = knots + (size_t)3 * nknots * i;
        double lo[3] = {K[0], K[1], K[2]}, hi[3] = {K[0], K[1], K[2]};
        for (int k = 1; k < nknots; ++k)
            for (int a_ = 0; a_ < 3; ++a_) { lo[a_] = std::min(lo[a_], K[3 * k + a_]); hi[a_] = std::max(hi[a_], K[3 * k + a_]); }
        double r2 = 0.0;
        for (int a_ = 0; a_ < 3; ++a_) c.kc[a_] = 0.5 * (lo[a_] + hi[a_]);
        for (int k = 0; k < nknots; ++k) {
            double d2 = 0.0;
            for (int a_ = 0; a_ < 3; ++a_) d2 += (K[3 * k + a_] - c.kc[a_]) * (K[3 * k + a_] - c.kc[a_]);
            r2 = std::max(r2, d2);
        }
        c.kr = std::sqrt(r2) * (1.0 + 1e-12);
    }
    if (kind == MODEL_THIN) return;
    int rB = (kind == MODEL_THERMAL || kind == MODEL_KNOT) ? 14 : 15;
    c.Bmag = std::exp(p[rB]);
    double st = std::sin(p[rB + 1]), ct = std::cos(p[rB + 1]), sp = std::sin(p[rB + 2]), cp = std::cos(p[rB + 2]);
    c.bh[0] = st * cp; c.bh[1] = st * sp; c.bh[2] = ct;
    sa.qn[0] = w; sa.qn[1] = x; sa.qn[2] = y; sa.qn[3] = z; sa.iqn = 1.0 / nq;
    sa.sthB = st; sa.cthB = ct; sa.sphB = sp; sa.cphB = cp;
    c.u[0] = p[rB + 3]; c.u[1] = p[rB + 4]; c.u[2] = p[rB + 5];
    if (kind == MODEL_THERMAL || kind == MODEL_KNOT) {
        double Th = std::exp(p[13]);
        double k0, k1;
        const double xb = 1.0 / Th;
        double k2 = host_besselk(xb, k0, k1);
        c.p1 = Th;
        c.p2 = k2 > 0.0 ? k0 / k2 : 1.0;
        c.p3 = k2 > 0.0 ? k1 / k2 : 1.0;
        // d/d ln Theta = -x d/dx at x = 1/Theta, with K0' = -K1, K1' = -K0 - K1/x, K2 = K0 + 2 K1/x
        if (k2 > 0.0) host_bessel_ratio_dlog(xb, k0, k1, k2, sa.dk0k2, sa.dk1k2);
    } else if (kind == MODEL_POWERLAW) {
        double pp = p[13];
        c.p1 = pp; c.p2 = std::exp(p[14]); c.p3 = gamma_max;
        c.g[0] = std::tgamma((3 * pp - 1) / 12.0) * std::tgamma((3 * pp + 19) / 12.0);
        c.g[1] = std::tgamma((3 * pp + 2) / 12.0) * std::tgamma((3 * pp + 22) / 12.0);
        // d/dp of the Gamma products: Gamma'(x) = Gamma(x) psi(x), dx/dp = 1/4
        sa.dpop[0] = c.g[0] * 0.25 * (host_digamma((3 * pp - 1) / 12.0) + host_digamma((3 * pp + 19) / 12.0));
        sa.dpop[1] = c.g[1] * 0.25 * (host_digamma((3 * pp + 2) / 12.0) + host_digamma((3 * pp + 22) / 12.0));
    } else {
        double kap = p[13], wk = std::exp(p[14]);
        // the 2F1 factor and its partials in kappa and ln w, by two-partial duals through its series
        DN<2, double> kh = DN<2, double>::seed(kap, 0), wh(wk);
        wh.d[1] = wk;
        const DN<2, double> hyp = host_kappa_hyp(kh, wh);
        c.p1 = kap; c.p2 = wk; c.p3 = hyp.v;
        c.g[0] = std::tgamma(kap - 4.0 / 3.0); c.g[1] = std::tgamma(kap - 2.0);
        c.g[2] = std::tgamma(kap / 4.0 - 1.0 / 3.0); c.g[3] = std::tgamma(kap / 4.0 + 4.0 / 3.0);
        c.g[4] = std::tgamma(2.0 + kap / 2.0);
        double k0, k1;
        const double xw = 1.0 / wk;
        double k2 = host_besselk(xw, k0, k1);
        c.g[5] = k0 / (k2 + 1e-40);
        sa.dpop[0] = hyp.d[0];
        sa.dpop[1] = hyp.d[1];
        sa.dpop[2] = c.g[0] * host_digamma(kap - 4.0 / 3.0);
        sa.dpop[3] = c.g[1] * host_digamma(kap - 2.0);
        sa.dpop[4] = c.g[2] * 0.25 * host_digamma(kap / 4.0 - 1.0 / 3.0);
        sa.dpop[5] = c.g[3] * 0.25 * host_digamma(kap / 4.0 + 4.0 / 3.0);
        sa.dpop[6] = c.g[4] * 0.5 * host_digamma(2.0 + kap / 2.0);
        // d(K0 / (K2 + 1e-40))/d ln w at 1/w (the 1e-40 guard is far below K2 for every w of the fits)
        double dk0k2 = 0.0, dk1k2 = 0.0;
        if (k2 > 0.0) host_bessel_ratio_dlog(xw, k0, k1, k2, dk0k2, dk1k2);
        sa.dpop[7] = dk0k2;
        // the fits' per-splat factors with their partials in (kappa, ln w, ln B), by three-partial duals
        using D3 = DN<3, double>;
        const D3 kd = D3::seed(kap, 0);
        D3 wd(wk); wd.d[1] = wk;
        D3 hd(c.p3); hd.d[0] = sa.dpop[0]; hd.d[1] = sa.dpop[1];
        D3 gd[6];
        for (int k = 0; k < 5; ++k) { gd[k] = D3(c.g[k]); gd[k].d[0] = sa.dpop[2 + k]; }
        gd[5] = D3(c.g[5]); gd[5].d[1] = sa.dpop[7];
        D3 Bd(c.Bmag); Bd.d[2] = c.Bmag;
        KappaK<D3> Kd;
        kappa_consts(kd, wd, hd, gd, Bd, Kd);
        for (int f = 0; f < KK_N; ++f) {
            c.kk.v[f] = Kd.v[f].v;
            for (int q = 0; q < 3; ++q) sa.dkk[q][f] = Kd.v[f].d[q];
        }
    }
}

}  // namespace ksd
