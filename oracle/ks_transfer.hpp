// Polarized synchrotron radiative transfer for the CPU oracle (test infrastructure; see
// ks_scalar.hpp). Restates src/Transfer/*.jl of the reference:
//   * cgs constants of ipole/symphony (constants.jl:3-12);
//   * thermal (Dexter 2016 emissivities, Kirchhoff absorptivities, Dexter rho_Q, Shcherbakov 2008
//     rho_V; Pandya+ 2016 and Leung+ 2011 alternatives), power-law (Pandya+ 2016, Jones & O'Dell
//     rotativities) and kappa (Pandya+ 2016, Marszewski+ 2021) fits (coefficients.jl, kappa.jl);
//     the fitting formulae are published and are written out as such;
//   * the polarization cap, the nu-hat invariants and the rotation into the screen basis
//     (coefficients.jl:264-314, step.jl:407-413);
//   * the exact constant-coefficient 4x4 step O = exp(-K D), Ej = int_0^D exp(-K u) du
//     (step.jl:27-243), here from the spectral projectors of K'^2 (eigenvalues L1^2 and -L2^2);
//   * the fluid frame, pitch angle and Walker-Penrose screen angle (frames.jl:31-128), with the
//     Boyer-Lindquist metric and ZAMO tetrad (spatial axes r, phi, -theta) written out.
#pragma once
#include "ks_scalar.hpp"
#include "ks_elliptic.hpp"

namespace kso {

// ---- constants (cgs, ipole/symphony values) ----------------------------------------------------
constexpr double ME = 9.1093826e-28, MP = 1.67262171e-24, CL = 2.99792458e10, EE = 4.80320680e-10;
constexpr double HPL = 6.6260693e-27, KBOL = 1.3806505e-16, GNEWT = 6.6742e-8, MSUN = 1.989e33;
constexpr double PC = 3.085678e18, JY = 1e-23;
constexpr double NU0 = 1e11;
constexpr double PLANCK0 = 2.0 * HPL * NU0 * NU0 * NU0 / (CL * CL);
constexpr double WEIGHT_CUTOFF = 1e-6;
inline double support_radius2() { return -2.0 * std::log(WEIGHT_CUTOFF); }
inline double gravitational_radius(double msun) { return GNEWT * msun * MSUN / (CL * CL); }

template <class T> struct Coef { T jI, jQ, jV, aI, aQ, aV, rQ, rV; };
template <class T> Coef<T> coef_zero() { return {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

// ---- modified Bessel functions K0, K1 -----------------------------------------------------------
// K_nu(x) = int_0^inf exp(-x cosh s) cosh(nu s) ds by the trapezoidal rule, which converges
// exponentially for this entire integrand (error ~ exp(-pi^2/h)); h = 1/8 gives ~1e-34.
template <class T> void besselk01(T x, T& k0, T& k1) {
    const double h = 0.125;
    T s0 = 0.5 * kexp(-x), s1 = 0.0;   // s = 0 endpoint: weight 1/2, cosh(0) = 1 (K1's term is 1/2 e^{-x})
    s1 = s0;
    for (int i = 1; i < 4000; ++i) {
        double s = h * i;
        T e = kexp(-x * std::cosh(s));
        s0 = s0 + e;
        s1 = s1 + e * std::cosh(s);
        if (val(e) < 1e-40 * val(s0)) break;
    }
    k0 = h * s0; k1 = h * s1;
}
template <class T> void besselk012(T x, T& k0, T& k1, T& k2) {
    besselk01(x, k0, k1);
    k2 = k0 + 2.0 * k1 / x;
}

// ---- Planck function in the nu-hat scaling (coefficients.jl:45-67) -------------------------------
template <class T> T planck(T nu, T Th) {
    T x = HPL * nu / (ME * CL * CL * Th);
    T nh = nu / NU0;
    T pref = PLANCK0 * nh * nh * nh;
    if (x < 2e-3) return pref / (x / 24.0 * (24.0 + x * (12.0 + x * (4.0 + x))));
    return pref / (kexp(x) - 1.0);
}

// ---- thermal synchrotron (coefficients.jl:86-151) ---------------------------------------------------
template <class T> T dexter_II(T x) { return 2.5651 * (1.0 + 1.92 * kpow(x, -1.0 / 3.0) + 0.9977 * kpow(x, -2.0 / 3.0)) * kexp(-1.8899 * kpow(x, 1.0 / 3.0)); }
template <class T> T dexter_IQ(T x) { return 2.5651 * (1.0 + 0.93193 * kpow(x, -1.0 / 3.0) + 0.499873 * kpow(x, -2.0 / 3.0)) * kexp(-1.8899 * kpow(x, 1.0 / 3.0)); }
template <class T> T dexter_IV(T x) {
    return (1.81384 / x + 3.42319 * kpow(x, -2.0 / 3.0) + 0.0292545 * kpow(x, -0.5) + 2.03773 * kpow(x, -1.0 / 3.0)) * kexp(-1.8899 * kpow(x, 1.0 / 3.0));
}

// Pandya+ (2016) thermal emissivities (coefficients.jl:153-175)
template <class T> void thermal_pandya(T ne, T Th, T B, T nu, T th, T& jI, T& jQ, T& jV) {
    T sth = ksin(th);
    T nuc = EE * B / (2.0 * KS_PI * ME * CL);
    T nus = (2.0 / 9.0) * nuc * sth * Th * Th;
    T X = nu / nus;
    T pref = ne * EE * EE * nuc / CL;
    T term1 = std::sqrt(2.0) * KS_PI / 27.0 * sth;
    T e13 = kexp(-kpow(X, 1.0 / 3.0));
    T a = kpow(X, 0.5) + std::pow(2.0, 11.0 / 12.0) * kpow(X, 1.0 / 6.0);
    jI = pref * term1 * a * a * e13;
    T Tp = kpow(Th, 24.0 / 25.0);
    T t2 = (7.0 * Tp + 35.0) / (10.0 * Tp + 75.0);
    T b = kpow(X, 0.5) + t2 * std::pow(2.0, 11.0 / 12.0) * kpow(X, 1.0 / 6.0);
    jQ = pref * term1 * b * b * e13;
    T tv1 = (37.0 - 87.0 * ksin(th - 28.0 / 25.0)) / (100.0 * (Th + 1.0));
    T tv2 = kpow(1.0 + (kpow(Th, 3.0 / 5.0) / 25.0 + 7.0 / 10.0) * kpow(X, 9.0 / 25.0), 5.0 / 3.0);
    jV = pref * tv1 * tv2 * e13;
}

// Leung+ (2011) unpolarized emissivity (coefficients.jl:177-189)
template <class T> T thermal_leung(T ne, T Th, T B, T nu, T th) {
    T k0, k1, k2;
    besselk012(1.0 / Th, k0, k1, k2);
    T K2 = kmax(k2, T(1e-40));
    T nuc = EE * B / (2.0 * KS_PI * ME * CL);
    T nus = (2.0 / 9.0) * nuc * Th * Th * ksin(th);
    if (nu > 1e12 * nus) return 0.0;
    T x = nu / nus;
    T a = kpow(x, 0.5) + std::pow(2.0, 11.0 / 12.0) * kpow(x, 1.0 / 6.0);
    return (std::sqrt(2.0) * KS_PI * EE * EE * ne * nus / (3.0 * CL * K2)) * a * a * kexp(-kpow(x, 1.0 / 3.0));
}

// The thermal coefficients given the Bessel values at 1/Theta_e (k0, k1, k2), which depend only on
// the temperature (per splat, constant along the ray).
template <class T> Coef<T> thermal_synchrotron_k(T ne, T Th, T B, T nu, T th, T k0, T k1, T k2, bool dexter_rhoV = false, bool pandya = false) {
    T sth = ksin(th), cth = kcos(th);
    T w0 = EE * B / (ME * CL);
    T wp2 = 4.0 * KS_PI * ne * EE * EE / ME;
    T w = 2.0 * KS_PI * nu;
    T w0w = w0 / w;
    T wp2w2 = wp2 / (w * w);
    if (!(sth > 0.0)) {
        T f0 = k2 > 0.0 ? k0 / k2 : T(1.0);
        Coef<T> c = coef_zero<T>();
        c.rV = w / CL * (wp2w2 * w0w * f0 * cth);
        return c;
    }
    T x = Th * ksqrt(std::sqrt(2.0) * sth * (1e3 * w0w));
    T extraterm = (0.011 * kexp(-x / 47.2) - std::pow(2.0, -1.0 / 3.0) / std::pow(3.0, 23.0 / 6.0) * KS_PI * 1e4 * kpow(x + 1e-16, -8.0 / 3.0)) *
                  (0.5 + 0.5 * ktanh((klog(x) - std::log(120.0)) / 0.1));
    T jffunc = 2.011 * kexp(-kpow(x, 1.035) / 4.7) - kcos(x / 2.0) * kexp(-kpow(x, 1.2) / 2.73) - 0.011 * kexp(-x / 47.2) + extraterm;
    T kratio = k2 > 0.0 ? k1 / k2 : T(1.0);
    T eps11m22 = jffunc * wp2w2 * w0w * w0w * (kratio + 6.0 * Th) * sth * sth;
    T rQ = w / (2.0 * CL) * eps11m22;
    T ff;
    if (dexter_rhoV && k2 > 0.0) ff = (k0 - 0.43793091 * klog(1.0 + 0.00185777 * kpow(x, 1.50316886))) / k2;
    else ff = (k2 > 0.0 ? k0 / k2 : T(1.0)) * (1.0 - 0.11 * klog(1.0 + 0.035 * x));
    T rV = w / CL * (wp2w2 * w0w * ff * cth);
    T jI, jQ, jV;
    if (pandya) {
        thermal_pandya(ne, Th, B, nu, th, jI, jQ, jV);
    } else {
        T nus = 3.0 * EE * B * sth / (4.0 * KS_PI * ME * CL) * Th * Th + 1.0;
        T xs = nu / nus;
        T pref = ne * EE * EE * nu / (2.0 * std::sqrt(3.0) * CL * Th * Th);
        jI = pref * dexter_II(xs);
        jQ = pref * dexter_IQ(xs);
        jV = 2.0 * ne * EE * EE * nu * cth / sth / (3.0 * std::sqrt(3.0) * CL * Th * Th * Th) * dexter_IV(xs);
    }
    T Bn = planck(nu, Th);
    T ib = Bn > 0.0 ? 1.0 / Bn : T(0.0);
    return {jI, jQ, jV, jI * ib, jQ * ib, jV * ib, rQ, rV};
}
template <class T> Coef<T> thermal_synchrotron(T ne, T Th, T B, T nu, T th, bool dexter_rhoV = false, bool pandya = false) {
    T k0, k1, k2;
    besselk012(1.0 / Th, k0, k1, k2);
    return thermal_synchrotron_k(ne, Th, B, nu, th, k0, k1, k2, dexter_rhoV, pandya);
}

// ---- power law (coefficients.jl:208-256) -------------------------------------------------------------------
template <class T> void powerlaw_rotativities(T ne, T p, T gmin, T gmax, T B, T nu, T th, T& rQ, T& rV) {
    T sth = ksin(th), cth = kcos(th);
    T nuB = EE * B * sth / (2.0 * KS_PI * ME * CL);
    T rperp = ne * EE * EE * (p - 1.0) / (ME * CL * nuB * (kpow(gmin, 1.0 - p) - kpow(gmax, 1.0 - p)));
    T numin = gmin * gmin * nuB;
    rQ = -rperp * kpow(nuB / nu, T(3.0)) * kpow(gmin, 2.0 - p) * (1.0 - kpow(numin / nu, p / 2.0 - 1.0)) / (p / 2.0 - 1.0);
    rV = 2.0 * (p + 2.0) / (p + 1.0) * rperp * kpow(nuB / nu, T(2.0)) * kpow(gmin, -(p + 1.0)) * klog(gmin) * cth / sth;
}
template <class T> Coef<T> powerlaw_synchrotron(T ne, T p, T gmin, T gmax, T B, T nu, T th, bool rotativities = true) {
    T sth = ksin(th), cth = kcos(th);
    T nuc = EE * B / (2.0 * KS_PI * ME * CL);
    T gspan = kpow(gmin, 1.0 - p) - kpow(gmax, 1.0 - p);
    T pref = ne * EE * EE * nuc / CL;
    T jI = pref * (kpow(T(3.0), p / 2.0) * (p - 1.0) * sth) / (2.0 * (p + 1.0) * gspan) *
           kgamma((3.0 * p - 1.0) / 12.0) * kgamma((3.0 * p + 19.0) / 12.0) * kpow(nu / (nuc * sth), -(p - 1.0) / 2.0);
    T jQ = (p + 1.0) / (p + 7.0 / 3.0) * jI;
    T jV = (171.0 / 250.0) * kpow(p, T(49.0 / 100.0)) * cth / sth * kpow(nu / (3.0 * nuc * sth), T(-0.5)) * jI;
    T apref = ne * EE * EE / (nu * ME * CL);
    T aI = apref * kpow(T(3.0), (p + 1.0) / 2.0) * (p - 1.0) / (4.0 * gspan) *
           kgamma((3.0 * p + 2.0) / 12.0) * kgamma((3.0 * p + 22.0) / 12.0) * kpow(nu / (nuc * sth), -(p + 2.0) / 2.0);
    T aQ = aI * kpow((17.0 / 500.0) * p - 43.0 / 1250.0, T(43.0 / 500.0));
    T term6 = kpow((31.0 / 10.0) * kpow(sth, T(-48.0 / 25.0)) - 31.0 / 10.0, T(64.0 / 125.0));
    T sgn = cth > 0.0 ? T(1.0) : (cth < 0.0 ? T(-1.0) : T(0.0));
    T aV = aI * kpow((71.0 / 100.0) * p + 22.0 / 625.0, T(197.0 / 500.0)) * term6 * kpow(nu / (nuc * sth), T(-0.5)) * sgn;
    T rQ = 0.0, rV = 0.0;
    if (rotativities) powerlaw_rotativities(ne, p, gmin, gmax, B, nu, th, rQ, rV);
    return {jI, jQ, jV, aI, aQ, aV, rQ, rV};
}

// ---- kappa (kappa.jl) -----------------------------------------------------------------------------------------
// Gauss hypergeometric 2F1(a, b; c; z) for z <= 0 through the Pfaff transformation
// 2F1(a, b; c; z) = (1 - z)^(-a) 2F1(a, c - b; c; z/(z - 1)), whose series converges for z < 0.
inline double hyp2f1_neg(double a, double b, double c, double z) {
    double w = z / (z - 1.0);
    double bb = c - b, term = 1.0, sum = 1.0;
    for (int n = 0; n < 200000; ++n) {
        term *= (a + n) * (bb + n) / ((c + n) * (n + 1.0)) * w;
        sum += term;
        if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
    }
    return std::pow(1.0 - z, -a) * sum;
}
inline double kappa_hypergeometric(double kappa, double w) { return hyp2f1_neg(kappa - 1.0 / 3.0, kappa + 1.0, kappa + 2.0 / 3.0, -kappa * w); }
// the same series in any scalar (dual numbers carry the parameters' and the argument's derivatives)
template <class T> T hyp2f1_neg_t(T a, T b, T c, T z) {
    T w = z / (z - 1.0);
    T bb = c - b, term = T(1.0), sum = T(1.0);
    for (int n = 0; n < 200000; ++n) {
        term = term * (a + double(n)) * (bb + double(n)) / ((c + double(n)) * (n + 1.0)) * w;
        sum = sum + term;
        if (std::fabs(val(term)) < 1e-17 * std::fabs(val(sum))) break;
    }
    return kpow(1.0 - z, -a) * sum;
}
template <class T> T kappa_hypergeometric_t(T kappa, T w) { return hyp2f1_neg_t(kappa - 1.0 / 3.0, kappa + 1.0, kappa + 2.0 / 3.0, -kappa * w); }

template <class T> T kappa_rhoQ_fit(int i, T w, T X) {
    T sw = ksqrt(w), e5 = kexp(-5.0 * w);
    if (i == 1) return (17.0 * w - 3.0 * sw + 7.0 * sw * e5) * (1.0 - kexp(-kpow(X, T(0.84)) / 30.0) - ksin(X / 10.0) * kexp(-3.0 * kpow(X, T(0.471)) / 2.0));
    if (i == 2) return (46.0 * w / 3.0 - 5.0 * sw / 3.0 + 17.0 * sw / 3.0 * e5) * (1.0 - kexp(-kpow(X, T(0.84)) / 18.0) - ksin(X / 6.0) * kexp(-7.0 * kpow(X, T(0.5)) / 4.0));
    if (i == 3) return (14.0 * w - 13.0 * sw / 8.0 + 9.0 * sw / 2.0 * e5) * (1.0 - kexp(-kpow(X, T(0.84)) / 12.0) - ksin(X / 4.0) * kexp(-2.0 * kpow(X, T(0.525))));
    return (25.0 * w / 2.0 - sw + 5.0 * sw * e5) * (1.0 - kexp(-kpow(X, T(0.84)) / 8.0) - ksin(3.0 * X / 8.0) * kexp(-9.0 * kpow(X, T(0.541)) / 4.0));
}
template <class T> T kappa_rhoV_fit(int i, T w, T X) {
    if (i == 1) return (w * w + 2.0 * w + 1.0) / (25.0 * w * w / 8.0 + 4.0 * w + 1.0) * (1.0 - 0.17 * klog(1.0 + 0.447 * kpow(X, T(-0.5))));
    if (i == 2) return (w * w + 54.0 * w + 50.0) / (30.0 * w * w / 11.0 + 134.0 * w + 50.0) * (1.0 - 0.17 * klog(1.0 + 0.391 * kpow(X, T(-0.5))));
    if (i == 3) return (w * w + 43.0 * w + 38.0) / (7.0 * w * w / 3.0 + 185.0 * w / 2.0 + 38.0) * (1.0 - 0.17 * klog(1.0 + 0.348 * kpow(X, T(-0.5))));
    return (w + 13.0 / 14.0) / (2.0 * w + 13.0 / 14.0) * (1.0 - 0.17 * klog(1.0 + 0.313 * kpow(X, T(-0.5))));
}
template <class T> void kappa_rotativities(T ne, T kappa, T w, T B, T nu, T th, T& rQ, T& rV, T k0k2 = T(-1.0)) {
    T sth = ksin(th), cth = kcos(th);
    T nuc = EE * B / (2.0 * KS_PI * ME * CL);
    T X = nu / ((w * kappa) * (w * kappa) * nuc * sth);
    T kc = kmin(kmax(kappa, T(3.5)), T(5.0));
    int i = kc < 4.0 ? 1 : (kc < 4.5 ? 2 : 3);
    T klo = 3.5 + (i - 1) / 2.0;
    T f = (kc - klo) * 2.0;
    T fQ = (1.0 - f) * kappa_rhoQ_fit(i, w, X) + f * kappa_rhoQ_fit(i + 1, w, X);
    T fV = (1.0 - f) * kappa_rhoV_fit(i, w, X) + f * kappa_rhoV_fit(i + 1, w, X);
    if (k0k2 < 0.0) {
        T k0, k1, k2;
        besselk012(1.0 / w, k0, k1, k2);
        k0k2 = k0 / (k2 + 1e-40);
    }
    rQ = -(ne * EE * EE * nuc * nuc * sth * sth) / (ME * CL * nu * nu * nu) * fQ;
    rV = 2.0 * (ne * EE * EE * nuc * cth) / (ME * CL * nu * nu) * k0k2 * fV;
}
template <class T> Coef<T> kappa_synchrotron(T ne, T kappa, T w, T B, T nu, T th, T hyp, T k0k2 = T(-1.0)) {
    T sth = ksin(th), cth = kcos(th);
    T nuc = EE * B / (2.0 * KS_PI * ME * CL);
    T nuw = (w * kappa) * (w * kappa) * nuc * sth;
    T X = nu / nuw;
    const double small = 1e-40;
    T pref = ne * EE * EE * nuc * sth / CL;
    T g43 = kgamma(kappa - 4.0 / 3.0), g2 = kgamma(kappa - 2.0);
    T gq = kgamma(kappa / 4.0 - 1.0 / 3.0), gp = kgamma(kappa / 4.0 + 4.0 / 3.0);
    T c37 = std::pow(3.0, 7.0 / 3.0);
    T NlowI = 4.0 * KS_PI * g43 / (c37 * g2);
    T NhighI = 0.25 * kpow(T(3.0), (kappa - 1.0) / 2.0) * (kappa - 2.0) * (kappa - 1.0) * gq * gp + small;
    T xI = 3.0 * kpow(kappa, T(-1.5));
    T jI = pref * NlowI * kpow(X, T(1.0 / 3.0)) * kpow(1.0 + kpow(X, xI * (3.0 * kappa - 4.0) / 6.0) * kpow(NlowI / NhighI, xI), -1.0 / xI);
    T NlowQ = -0.5 * 4.0 * KS_PI * g43 / (c37 * g2);
    T NhighQ = -((4.0 / 5.0) * (4.0 / 5.0) + kappa / 50.0) * 0.25 * kpow(T(3.0), (kappa - 1.0) / 2.0) * (kappa - 2.0) * (kappa - 1.0) * gq * gp + small;
    T xQ = (37.0 / 10.0) * kpow(kappa, T(-8.0 / 5.0));
    T jQ = -(pref * NlowQ * kpow(X, T(1.0 / 3.0)) * kpow(1.0 + kpow(X, xQ * (3.0 * kappa - 4.0) / 6.0) * kpow(NlowQ / NhighQ, xQ), -1.0 / xQ));
    T NlowV = -(3.0 / 4.0) * (3.0 / 4.0) * kpow(kpow(sth, T(-12.0 / 5.0)) - 1.0, T(12.0 / 25.0)) * (kpow(kappa, T(-66.0 / 125.0)) / w) * kpow(X, T(-7.0 / 20.0)) *
              4.0 * KS_PI * g43 / (c37 * g2);
    T NhighV = -(7.0 / 8.0) * (7.0 / 8.0) * kpow(kpow(sth, T(-5.0 / 2.0)) - 1.0, T(11.0 / 25.0)) * (kpow(kappa, T(-11.0 / 25.0)) / w) * kpow(X, T(-0.5)) *
               0.25 * kpow(T(3.0), (kappa - 1.0) / 2.0) * (kappa - 2.0) * (kappa - 1.0) * gq * gp - small;
    T xV = 3.0 * kpow(kappa, T(-1.5));
    T sgn = cth > 0.0 ? T(1.0) : (cth < 0.0 ? T(-1.0) : T(0.0));
    T jV = val(kabs(NhighV)) < small * small ? T(0.0)
         : -(pref * NlowV * kpow(X, T(1.0 / 3.0)) * kpow(1.0 + kpow(X, xV * (3.0 * kappa - 4.0) / 6.0) * kpow(NlowV / NhighV, xV), -1.0 / xV)) * sgn;
    T apref = ne * EE / (B * sth);
    T g53 = T(std::tgamma(5.0 / 3.0)), gk2 = kgamma(2.0 + kappa / 2.0);
    T ck = (kappa - 2.0) * (kappa - 1.0) * kappa;
    T c16 = std::pow(3.0, 1.0 / 6.0);
    T tpi2 = (2.0 * KS_PI) * (2.0 * KS_PI);
    T wk = w * kappa;
    T NlowA = c16 * (10.0 / 41.0) * tpi2 / kpow(wk, 16.0 / 3.0 - kappa) * ck / (3.0 * kappa - 1.0) * g53 * hyp;
    T NhighA = 2.0 * std::pow(KS_PI, 2.5) / 3.0 * ck / kpow(wk, T(5.0)) * (2.0 * gk2 / (2.0 + kappa) - 1.0) * (kpow(3.0 / kappa, T(19.0 / 4.0)) + 3.0 / 5.0) + small;
    T xA = kpow(-7.0 / 4.0 + 8.0 * kappa / 5.0, T(-43.0 / 50.0));
    T aI = apref * NlowA * kpow(X, T(-5.0 / 3.0)) * kpow(1.0 + kpow(X, xA * (3.0 * kappa - 1.0) / 6.0) * kpow(NlowA / NhighA, xA), -1.0 / xA);
    T NlowQA = -(25.0 / 48.0) * c16 * (10.0 / 41.0) * tpi2 / kpow(wk, 16.0 / 3.0 - kappa) * ck / (3.0 * kappa - 1.0) * g53 * hyp;
    T NhighQA = -(21.0 * 21.0 * kpow(kappa, T(-144.0 / 25.0)) + 11.0 / 20.0) * 2.0 * std::pow(KS_PI, 2.5) / 3.0 * ck / kpow(wk, T(5.0)) * (2.0 * gk2 / (2.0 + kappa) - 1.0) + small;
    T xQA = (7.0 / 5.0) * kpow(kappa, T(-23.0 / 20.0));
    T aQ = -(apref * NlowQA * kpow(X, T(-5.0 / 3.0)) * kpow(1.0 + kpow(X, xQA * (3.0 * kappa - 1.0) / 6.0) * kpow(NlowQA / NhighQA, xQA), -1.0 / xQA));
    T NlowVA = -(77.0 / (100.0 * w)) * kpow(kpow(sth, T(-114.0 / 50.0)) - 1.0, T(223.0 / 500.0)) * kpow(X, T(-7.0 / 20.0)) * kpow(kappa, T(-7.0 / 10.0)) *
               c16 * (10.0 / 41.0) * tpi2 / kpow(wk, 16.0 / 3.0 - kappa) * ck / (3.0 * kappa - 1.0) * g53 * hyp;
    T NhighVA = -(143.0 / 10.0 * kpow(w, T(-116.0 / 125.0))) * kpow(kpow(sth, T(-41.0 / 20.0)) - 1.0, T(0.5)) *
                (169.0 * kpow(kappa, T(-8.0)) + 13.0 / 2500.0 * kappa - 263.0 / 5000.0 + 47.0 / (200.0 * kappa)) * kpow(X, T(-0.5)) *
                2.0 * std::pow(KS_PI, 2.5) / 3.0 * ck / kpow(wk, T(5.0)) * (2.0 * gk2 / (2.0 + kappa) - 1.0) + small;
    T xVA = (61.0 / 50.0) * kpow(kappa, T(-142.0 / 125.0)) + 7.0 / 1000.0;
    T aV = -(apref * NlowVA * kpow(X, T(-5.0 / 3.0)) * kpow(1.0 + kpow(X, xVA * (3.0 * kappa - 1.0) / 6.0) * kpow(NlowVA / NhighVA, xVA), -1.0 / xVA)) * sgn;
    T rQ, rV;
    kappa_rotativities(ne, kappa, w, B, nu, th, rQ, rV, k0k2);
    return {jI, jQ, jV, aI, aQ, aV, rQ, rV};
}

// ---- cap, invariants, screen rotation --------------------------------------------------------------------------
template <class T> T pow2_below(T x) {
    double v = val(x);
    if (!(std::isfinite(v) && v > 0)) return T(1.0);
    int e; std::frexp(v, &e);
    return T(std::ldexp(1.0, e - 1));
}
template <class T> T cap_factor(T I, T Q, T V, double fmax) {
    if (!(I > 0.0)) return (val(Q) == 0.0 && val(V) == 0.0) ? T(1.0) : T(0.0);
    T s = pow2_below(kmax(kmax(kabs(I), kabs(Q)), kabs(V)));
    T Ih = I / s, q = Q / s, v = V / s;
    T pp = q * q + v * v;
    T p = pp > 0.0 ? ksqrt(pp) : T(0.0);
    return p > fmax * Ih ? fmax * Ih / p : T(1.0);
}
template <class T> Coef<T> cap_polarization(const Coef<T>& c, double fmax = 0.99) {
    T fe = cap_factor(c.jI, c.jQ, c.jV, fmax), fa = cap_factor(c.aI, c.aQ, c.aV, fmax);
    return {c.jI, fe * c.jQ, fe * c.jV, c.aI, fa * c.aQ, fa * c.aV, c.rQ, c.rV};
}
template <class T> Coef<T> invariants(const Coef<T>& c, T nu) {
    T nh = nu / NU0;
    T i2 = 1.0 / (nh * nh);
    return {c.jI * i2, c.jQ * i2, c.jV * i2, nh * c.aI, nh * c.aQ, nh * c.aV, nh * c.rQ, nh * c.rV};
}
template <class T> struct Screen { T j[4], a[4], r[3]; };
template <class T> Screen<T> rotate_to_screen(const Coef<T>& c, T chi) {
    T s2 = ksin(2.0 * chi), c2 = kcos(2.0 * chi);
    Screen<T> s;
    s.j[0] = c.jI; s.j[1] = c.jQ * c2; s.j[2] = c.jQ * s2; s.j[3] = c.jV;
    s.a[0] = c.aI; s.a[1] = c.aQ * c2; s.a[2] = c.aQ * s2; s.a[3] = c.aV;
    s.r[0] = c.rQ * c2; s.r[1] = c.rQ * s2; s.r[2] = c.rV;
    return s;
}

// ---- the exact 4x4 step ----------------------------------------------------------------------------------------------
template <class T> struct M4 { T m[4][4]; };
template <class T> M4<T> m4_zero() { M4<T> r; for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.m[i][j] = 0.0; return r; }
template <class T> M4<T> m4_eye() { M4<T> r = m4_zero<T>(); for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0; return r; }
template <class T> M4<T> m4_mul(const M4<T>& A, const M4<T>& B) {
    M4<T> r;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) {
        T s = A.m[i][0] * B.m[0][j];
        for (int k = 1; k < 4; ++k) s = s + A.m[i][k] * B.m[k][j];
        r.m[i][j] = s;
    }
    return r;
}
// r = c0 I + c1 K + c2 K2 + c3 K3
template <class T> M4<T> m4_comb(T c0, T c1, const M4<T>& K, T c2, const M4<T>& K2, T c3, const M4<T>& K3) {
    M4<T> r;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j)
        r.m[i][j] = c1 * K.m[i][j] + c2 * K2.m[i][j] + c3 * K3.m[i][j] + (i == j ? c0 : T(0.0));
    return r;
}

// e(a) = (1 - e^{-a})/a = int_0^1 e^{-at} dt
template <class T> T phi1(T a) {
    if (val(kabs(a)) < 1e-3) return 1.0 - a / 2.0 + a * a / 6.0 - a * a * a / 24.0 + a * a * a * a / 120.0;
    return -kexpm1(-a) / a;
}
// M_n(a) = int_0^1 t^n e^{-at} dt, n = 0..9
// Power-series coefficients of the moments, c[n][m] = (-1)^m / (m! (n + m + 1)), m = 0..23
struct MomentSeries {
    double c[10][24];
    constexpr MomentSeries() : c{} {
        for (int n = 0; n < 10; ++n) {
            double fact = 1.0;
            for (int m = 0; m < 24; ++m) {
                if (m > 0) fact *= m;
                c[n][m] = ((m % 2) ? -1.0 : 1.0) / (fact * (n + m + 1));
            }
        }
    }
};
inline constexpr MomentSeries MOMENT_SERIES{};

// M_n(a) = int_0^1 t^n e^{-at} dt for n = 0..nmax (step.jl:239-268): for |a| < 1 the 24-term power
// series by Horner's rule, otherwise the upward recurrence M_0 = (1 - e^{-a})/a,
// M_n = (n M_{n-1} - e^{-a})/a.
template <class T> void moments(T a, T M[10], int nmax = 9) {
    if (val(kabs(a)) < 1.0) {
        for (int n = 0; n <= nmax; ++n) {
            T s = T(MOMENT_SERIES.c[n][23]);
            for (int m = 22; m >= 0; --m) s = s * a + MOMENT_SERIES.c[n][m];
            M[n] = s;
        }
    } else {
        T e = kexp(-a);
        M[0] = (1.0 - e) / a;
        for (int n = 1; n <= nmax; ++n) M[n] = ((double)n * M[n - 1] - e) / a;
    }
}
// the integrals of the projector form given the moments M_0..M_9 when b < 0.1 (M_1 otherwise)
// int_0^1 e^{-at}(cosh bt - 1) dt
template <class T> T int_coshm1(T a, T b, const T M[10]) {
    if (b < 0.1) {
        T b2 = b * b;
        return b2 * (M[2] / 2.0 + b2 * (M[4] / 24.0 + b2 * (M[6] / 720.0 + b2 * M[8] / 40320.0)));
    }
    return (phi1(a - b) + phi1(a + b)) / 2.0 - phi1(a);
}
// int_0^1 e^{-at}(1 - cos bt) dt
template <class T> T int_versin(T a, T b, const T M[10]) {
    if (b < 0.1) {
        T b2 = b * b;
        return b2 * (M[2] / 2.0 - b2 * (M[4] / 24.0 - b2 * (M[6] / 720.0 - b2 * M[8] / 40320.0)));
    }
    T s = ksin(b), c = kcos(b);
    return phi1(a) - (a - kexp(-a) * (a * c - b * s)) / (a * a + b * b);
}
// int_0^1 e^{-at} cos(bt) dt
template <class T> T int_cos(T a, T b, const T M[10]) { return phi1(a) - int_versin(a, b, M); }
// int_0^1 e^{-at} sin(bt)/b dt
template <class T> T int_sinc(T a, T b, const T M[10]) {
    if (b < 0.1) {
        T b2 = b * b;
        return M[1] - b2 * (M[3] / 6.0 - b2 * (M[5] / 120.0 - b2 * (M[7] / 5040.0 - b2 * M[9] / 362880.0)));
    }
    T s = ksin(b), c = kcos(b);
    return (1.0 - kexp(-a) * (a * s / b + c)) / (a * a + b * b);
}
// int_0^1 e^{-at}(sinh(bt)/b - t) dt
template <class T> T int_sinhc_excess(T a, T b, const T M[10]) {
    if (b < 0.1) {
        T b2 = b * b;
        return b2 * (M[3] / 6.0 + b2 * (M[5] / 120.0 + b2 * (M[7] / 5040.0 + b2 * M[9] / 362880.0)));
    }
    return (phi1(a - b) - phi1(a + b)) / (2.0 * b) - M[1];
}
template <class T> T sinhc(T b) { return val(kabs(b)) < 0.1 ? 1.0 + b * b / 6.0 * (1.0 + b * b / 20.0 * (1.0 + b * b / 42.0 * (1.0 + b * b / 72.0))) : ksinh(b) / b; }
template <class T> T sinc_(T b) { return val(kabs(b)) < 0.1 ? 1.0 - b * b / 6.0 * (1.0 - b * b / 20.0 * (1.0 - b * b / 42.0 * (1.0 - b * b / 72.0))) : ksin(b) / b; }

// O = exp(-K D), Ej = int_0^D exp(-K u) du for K = alpha_I + K', K' as in step.jl:190
// The step as polynomials in K' (K = alpha_I + K'), the coefficients of step.jl:164-243:
// O = o0 + o1 K' + o2 K'^2 + o3 K'^3, Ej = Dp (e0 + e1 K' + e2 K'^2 + e3 K'^3), with
// K6 = (aQ, aU, aV, rQ, rU, rV) Dp the entries of K' after the power-of-two rescaling of the length.
// the coefficients from the invariants of the scaled K = alpha_I + K': aI (scaled), a2 = |(aQ, aU, aV)|^2,
// r2 = |(rQ, rU, rV)|^2, ar = (aQ, aU, aV).(rQ, rU, rV), and f = D / Dp (all the step depends on)
template <class T> void step_coefficients_inv(T aI, T a2, T r2, T ar, T f, T o[4], T e4[4]) {
    T a = aI * f;
    T e = kexp(-a);
    if ((a2 + r2) * f * f < 1e-200) {
        T M[10]; moments(a, M, 1);
        o[0] = e; o[1] = -f * e; o[2] = T(0.0); o[3] = T(0.0);
        e4[0] = f * phi1(a); e4[1] = -f * f * M[1]; e4[2] = T(0.0); e4[3] = T(0.0);
        return;
    }
    T h = (a2 - r2) / 2.0;
    T root = ksqrt(h * h + ar * ar);
    T L1sq, L2sq;
    if (h >= 0.0) { L1sq = h + root; L2sq = L1sq > 0.0 ? ar * ar / L1sq : T(0.0); }
    else { L2sq = root - h; L1sq = L2sq > 0.0 ? ar * ar / L2sq : T(0.0); }
    T L1 = L1sq > 0.0 ? ksqrt(L1sq) : T(0.0), L2 = L2sq > 0.0 ? ksqrt(L2sq) : T(0.0);
    T b1 = L1 * f, b2 = L2 * f;
    T Th = L1sq + L2sq;
    if (Th * f * f * (1.0 + (a2 + r2) * f * f) < 1e-17) {
        T M[10]; moments(a, M, 3);
        o[0] = e; o[1] = -f * e; o[2] = f * f / 2.0 * e; o[3] = -f * f * f / 6.0 * e;
        e4[0] = f * M[0]; e4[1] = -f * f * M[1]; e4[2] = f * f * f * M[2] / 2.0; e4[3] = -f * f * f * f * M[3] / 6.0;
        return;
    }
    // projector form: exp(-K'f) = [(K'^2 + L2^2) (cosh b1 - K' sinh(b1)/L1) + (L1^2 - K'^2)(cos b2 - K' sin(b2)/L2)]/Th
    T ech, esh, echm1;   // e^{-a} cosh b1, e^{-a} sinh(b1)/b1, e^{-a}(cosh b1 - 1)
    if (b1 < 1.0) {
        ech = e * kcosh(b1); esh = e * sinhc(b1); T sh = ksinh(b1 / 2.0); echm1 = e * 2.0 * sh * sh;
    } else {
        T em = kexp(-(a - b1)), ep = kexp(-(a + b1));
        ech = (em + ep) / 2.0; esh = (em - ep) / (2.0 * b1); echm1 = (em - 2.0 * e + ep) / 2.0;
    }
    T cb2 = kcos(b2), sb2h = ksin(b2 / 2.0);
    T versin2 = 2.0 * sb2h * sb2h;
    // O = c0 1 + c2 K'^2 - f esh M3 - f e sinc(b2) M2, M3 = (L2^2 K' + K'^3)/Th, M2 = (L1^2 K' - K'^3)/Th
    T s3 = f * esh, s2 = f * e * sinc_(b2);
    o[0] = (L2sq * ech + L1sq * e * cb2) / Th;
    o[1] = -(s3 * L2sq + s2 * L1sq) / Th;
    o[2] = (echm1 + e * versin2) / Th;
    o[3] = -(s3 - s2) / Th;
    T pa = phi1(a);
    T M[10]; moments(a, M, (b1 < 0.1 || b2 < 0.1) ? 9 : 1);
    T Ich = f * (pa + int_coshm1(a, b1, M));
    T Ic = f * int_cos(a, b2, M);
    T Ish = f * f * (M[1] + int_sinhc_excess(a, b1, M));
    T Is = f * f * int_sinc(a, b2, M);
    e4[0] = (L2sq * Ich + L1sq * Ic) / Th;
    e4[1] = -(Ish * L2sq + Is * L1sq) / Th;
    e4[2] = f * (int_coshm1(a, b1, M) + int_versin(a, b2, M)) / Th;
    e4[3] = -(Ish - Is) / Th;
}
template <class T> void step_coefficients(const T al[4], const T rh[3], T D, T& Dp, T K6[6], T o[4], T e4[4]) {
    // exact power-of-two rescaling of the length (step.jl:164-173)
    Dp = pow2_below(D);
    T f = D / Dp;
    T aQ = al[1] * Dp, aU = al[2] * Dp, aV = al[3] * Dp;
    T rQ = rh[0] * Dp, rU = rh[1] * Dp, rV = rh[2] * Dp;
    K6[0] = aQ; K6[1] = aU; K6[2] = aV; K6[3] = rQ; K6[4] = rU; K6[5] = rV;
    T a2 = aQ * aQ + aU * aU + aV * aV, r2 = rQ * rQ + rU * rU + rV * rV;
    T ar = aQ * rQ + aU * rU + aV * rV;
    step_coefficients_inv(al[0] * Dp, a2, r2, ar, f, o, e4);
}

// y = K' v and y = K'^T v (K'6 as above), for any mix of scalar types
template <class T, class U, class R> void kprime_mul(const T K6[6], const U v[4], R y[4]) {
    y[0] = K6[0] * v[1] + K6[1] * v[2] + K6[2] * v[3];
    y[1] = K6[0] * v[0] + K6[5] * v[2] - K6[4] * v[3];
    y[2] = K6[1] * v[0] - K6[5] * v[1] + K6[3] * v[3];
    y[3] = K6[2] * v[0] + K6[4] * v[1] - K6[3] * v[2];
}
template <class T, class U, class R> void kprime_tmul(const T K6[6], const U v[4], R y[4]) {
    y[0] = K6[0] * v[1] + K6[1] * v[2] + K6[2] * v[3];
    y[1] = K6[0] * v[0] - K6[5] * v[2] + K6[4] * v[3];
    y[2] = K6[1] * v[0] + K6[5] * v[1] - K6[3] * v[3];
    y[3] = K6[2] * v[0] - K6[4] * v[1] + K6[3] * v[2];
}

// O = exp(-K D) and Ej = int_0^D exp(-K u) du (step.jl:164-243) assembled from the coefficients
template <class T> void step_operator(const T al[4], const T rh[3], T D, M4<T>& O, M4<T>& Ej) {
    T Dp, K6[6], o[4], e4[4];
    step_coefficients(al, rh, D, Dp, K6, o, e4);
    M4<T> K = m4_zero<T>();
    K.m[0][1] = K6[0]; K.m[0][2] = K6[1]; K.m[0][3] = K6[2];
    K.m[1][0] = K6[0]; K.m[1][2] = K6[5]; K.m[1][3] = -K6[4];
    K.m[2][0] = K6[1]; K.m[2][1] = -K6[5]; K.m[2][3] = K6[3];
    K.m[3][0] = K6[2]; K.m[3][1] = K6[4]; K.m[3][2] = -K6[3];
    M4<T> K2 = m4_mul(K, K), K3 = m4_mul(K2, K);
    O = m4_comb(o[0], o[1], K, o[2], K2, o[3], K3);
    Ej = m4_comb(e4[0], e4[1], K, e4[2], K2, e4[3], K3);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) Ej.m[i][j] = Dp * Ej.m[i][j];
}

// ---- frames (frames.jl) ------------------------------------------------------------------------------------------------
template <class T> struct Frame { T g, cosB, chi; };

// covariant BL photon momentum (-1, +-sqrt(R)/Delta, +-sqrt(Theta), lambda)
template <class T> void momentum_bl_d(T a, T r, T th, T eta, T lam, bool nur, bool nuth, T pd[4]) {
    T w = r * r + a * a - a * lam;
    T Dl = r * r - 2.0 * r + a * a;
    T R = w * w - Dl * (eta + (lam - a) * (lam - a));
    T c = kcos(th), s = ksin(th);
    T Th = eta + a * a * c * c - lam * lam * c * c / (s * s);
    // R below 1e-14 w^2 is rounding (a sample on the radial turning point, where R = 0): its square root would be
    // noise of order 1e-7 w, taken as 0
    T sR = R > 1e-14 * (w * w) ? ksqrt(R) : T(0.0), sT = Th > 0.0 ? ksqrt(Th) : T(0.0);
    pd[0] = -1.0; pd[1] = (nur ? T(1.0) : T(-1.0)) * sR / Dl; pd[2] = (nuth ? T(1.0) : T(-1.0)) * sT; pd[3] = lam;
}
// inverse BL metric times a covariant vector (t, r, theta, phi)
template <class T> void metric_uu_mul(T a, T r, T th, const T pd[4], T pu[4]) {
    T c = kcos(th), s = ksin(th);
    T S = r * r + a * a * c * c, Dl = r * r - 2.0 * r + a * a;
    T Xi = (r * r + a * a) * (r * r + a * a) - a * a * Dl * s * s;
    T gtt = -Xi / (Dl * S), gtp = -2.0 * a * r / (Dl * S), gpp = (Dl - a * a * s * s) / (Dl * S * s * s);
    pu[0] = gtt * pd[0] + gtp * pd[3];
    pu[1] = Dl / S * pd[1];
    pu[2] = pd[2] / S;
    pu[3] = gtp * pd[0] + gpp * pd[3];
}
// ZAMO tetrad: BL contravariant -> ZAMO components (t, r-hat, phi-hat, -theta-hat), and back
template <class T> struct Zamo { T lapse, omega, sgrr, sgthth, sgpp; };
template <class T> Zamo<T> zamo(T a, T r, T th) {
    T c = kcos(th), s = ksin(th);
    T S = r * r + a * a * c * c, Dl = r * r - 2.0 * r + a * a;
    T Xi = (r * r + a * a) * (r * r + a * a) - a * a * Dl * s * s;
    return {ksqrt(Dl * S / Xi), 2.0 * a * r / Xi, ksqrt(S / Dl), ksqrt(S), s * ksqrt(Xi / S)};
}
template <class T> void bl_to_zamo(const Zamo<T>& z, const T v[4], T o[4]) {
    o[0] = z.lapse * v[0];
    o[1] = z.sgrr * v[1];
    o[2] = z.sgpp * (v[3] - z.omega * v[0]);
    o[3] = -z.sgthth * v[2];
}
template <class T> void zamo_to_bl(const Zamo<T>& z, const T o[4], T v[4]) {
    v[0] = o[0] / z.lapse;
    v[1] = o[1] / z.sgrr;
    v[2] = -o[3] / z.sgthth;
    v[3] = o[2] / z.sgpp + z.omega * v[0];
}
// Lorentz boost ZAMO -> fluid for the spatial 4-velocity u = gamma beta (frames.jl:31-41)
template <class T> void boost(const T u[3], T L[4][4]) {
    T u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    T g = ksqrt(1.0 + u2);
    T q = 1.0 / (g + 1.0);
    L[0][0] = g;
    for (int i = 0; i < 3; ++i) { L[0][i + 1] = -u[i]; L[i + 1][0] = -u[i]; }
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) L[i + 1][j + 1] = (i == j ? T(1.0) : T(0.0)) + u[i] * u[j] * q;
}
template <class T> void mat4vec(const T L[4][4], const T v[4], T o[4]) {
    for (int i = 0; i < 4; ++i) o[i] = L[i][0] * v[0] + L[i][1] * v[1] + L[i][2] * v[2] + L[i][3] * v[3];
}

// local_frame (frames.jl:108-128)
template <class T> Frame<T> local_frame(T a, T r, T th, T eta, T lam, bool nur, bool nuth, T alpha, T beta, T thetao, const T u[3], const T B[3]) {
    T pd[4], pu[4], pz[4], pf[4];
    momentum_bl_d(a, r, th, eta, lam, nur, nuth, pd);
    metric_uu_mul(a, r, th, pd, pu);
    Zamo<T> z = zamo(a, r, th);
    bl_to_zamo(z, pu, pz);
    T L[4][4]; boost(u, L);
    mat4vec(L, pz, pf);
    T g = 1.0 / pf[0];
    T k[3] = {pf[1] * g, pf[2] * g, pf[3] * g};
    T Bn = ksqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    T Bh[3];
    if (Bn > 0.0) { Bh[0] = B[0] / Bn; Bh[1] = B[1] / Bn; Bh[2] = B[2] / Bn; }
    else { Bh[0] = 0.0; Bh[1] = 0.0; Bh[2] = 1.0; }
    T cosB = k[0] * Bh[0] + k[1] * Bh[1] + k[2] * Bh[2];
    T fv[3] = {k[1] * Bh[2] - k[2] * Bh[1], k[2] * Bh[0] - k[0] * Bh[2], k[0] * Bh[1] - k[1] * Bh[0]};
    T fn = ksqrt(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2]);
    if (!(fn > 0.0)) return {g, cosB, T(0.0)};
    T ff[4] = {T(0.0), fv[0] / fn, fv[1] / fn, fv[2] / fn};
    T um[3] = {-u[0], -u[1], -u[2]};
    T Li[4][4]; boost(um, Li);
    T fz[4], fb[4];
    mat4vec(Li, ff, fz);
    zamo_to_bl(z, fz, fb);
    // Walker-Penrose constant (frames.jl:62-70)
    T st = ksin(th), ct = kcos(th);
    T A = pu[0] * fb[1] - pu[1] * fb[0] + a * st * st * (pu[1] * fb[3] - pu[3] * fb[1]);
    T Bw = ((r * r + a * a) * (pu[3] * fb[2] - pu[2] * fb[3]) - a * (pu[0] * fb[2] - pu[2] * fb[0])) * st;
    T k1 = A * r - Bw * a * ct, k2 = -(A * a * ct + Bw * r);
    T mu = -(alpha + a * ksin(thetao));
    T ea = beta * k2 - mu * k1, eb = beta * k1 + mu * k2;
    return {g, cosB, katan2(-ea, eb)};
}

// redshift factor of a ZAMO emitter: g = alpha_lapse / (1 - omega lambda)
template <class T> T zamo_redshift(T a, T r, T th, T lam) {
    Zamo<T> z = zamo(a, r, th);
    return z.lapse / (1.0 - z.omega * lam);
}

}  // namespace kso
