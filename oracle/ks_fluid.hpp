// Fluid models for end-to-end pins of the oracle's transport (TEST INFRASTRUCTURE; nothing in the
// product uses them). They plug into the same samples, frames, coefficients, invariants and 4x4 step as
// the splat models, one fluid element per sample:
//   * ipole's example RIAF (the reference's test/riaf_model.jl:1-93, after ipole's model/riaf): a thick
//     disk with power-law density and temperature, a toroidal field, Keplerian rotation outside the ISCO
//     and the conserved-quantity plunge inside, thermal synchrotron with Pandya emissivities;
//   * the Gold et al. (2020) analytic torus, unpolarized (test/test_gold2020.jl:14-53): j = n (nu/nu_p)^-alpha,
//     alpha_nu = A n (nu/nu_p)^-(2.5 + alpha), the normal-observer velocity u_mu ~ (-1, 0, 0, l);
//   * shells r1 <= r <= r2 with given fluid-frame coefficients, a ZAMO velocity and a field
//     (test/test_polarized.jl:33-51), for the Faraday-screen rotation and additivity checks.
#pragma once
#include "ks_splats.hpp"
#include <vector>

namespace kso {

// prograde ISCO radius (Bardeen, Press & Teukolsky 1972)
inline double isco_radius(double a) {
    const double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    const double z2 = std::sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - std::copysign(std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)), a);
}

// Boyer-Lindquist metric, covariant and contravariant, index order (t, r, theta, phi)
inline void bl_metric_dd(double a, double r, double th, double g[4][4]) {
    const double s = std::sin(th), c = std::cos(th);
    const double S = r * r + a * a * c * c, Dl = r * r - 2.0 * r + a * a;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
    g[0][0] = -(1.0 - 2.0 * r / S);
    g[0][3] = g[3][0] = -2.0 * a * r * s * s / S;
    g[1][1] = S / Dl;
    g[2][2] = S;
    g[3][3] = (r * r + a * a + 2.0 * a * a * r * s * s / S) * s * s;
}
inline void bl_metric_uu(double a, double r, double th, double g[4][4]) {
    const double s = std::sin(th), c = std::cos(th);
    const double S = r * r + a * a * c * c, Dl = r * r - 2.0 * r + a * a;
    const double Xi = (r * r + a * a) * (r * r + a * a) - a * a * Dl * s * s;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
    g[0][0] = -Xi / (Dl * S);
    g[0][3] = g[3][0] = -2.0 * a * r / (Dl * S);
    g[1][1] = Dl / S;
    g[2][2] = 1.0 / S;
    g[3][3] = (Dl - a * a * s * s) / (Dl * S * s * s);
}
inline void mv4(const double g[4][4], const double v[4], double o[4]) {
    for (int i = 0; i < 4; ++i) o[i] = g[i][0] * v[0] + g[i][1] * v[1] + g[i][2] * v[2] + g[i][3] * v[3];
}
inline double dot4(const double g[4][4], const double u[4], const double v[4]) {
    double w[4]; mv4(g, v, w);
    return u[0] * w[0] + u[1] * w[1] + u[2] * w[2] + u[3] * w[3];
}

// ---- ipole's RIAF (riaf_model.jl:12-41; ipole model/riaf/example.par) ----------------------------------
struct RIAF { double nth0, Te0, disk_h, pow_nth, pow_T, Ne_unit, Te_unit, eps, r_isco, rout; bool pandya; };
inline RIAF riaf_example(double a, bool pandya) { return {1.0, 1.0, 0.1, -1.1, -0.84, 3e7, 3e11, 0.1, isco_radius(a), 100.0, pandya}; }

// the RIAF's fluid-frame coefficients and frame at one sample (riaf_model.jl:45-92); false outside the model
inline bool riaf_element(const RIAF& m, const RayCtx& rc, const Sample<double>& s, double nu_obs, Coef<double>& cf, Frame<double>& fr) {
    const double a = rc.a, r = s.r, th = s.th;
    const double rh = 1.0 + std::sqrt(1.0 - a * a);
    if (!(r > rh + 0.1 && r < m.rout)) return false;
    const double sth = std::sin(th), cth = std::cos(th);
    const double zc = r * cth, rcyl = r * sth;
    const double n = m.nth0 * std::exp(-zc * zc / 2.0 / rcyl / rcyl / m.disk_h / m.disk_h) * std::pow(r, m.pow_nth) * m.Ne_unit;
    const double Th = m.Te0 * std::pow(r, m.pow_T) * m.Te_unit * KBOL / (ME * CL * CL);
    double bmag = std::sqrt(8.0 * M_PI * m.eps * n * MP * CL * CL / 6.0 / r);
    if (bmag == 0.0) bmag = 1e-6;
    double gdd[4][4], guu[4][4];
    bl_metric_dd(a, r, th, gdd);
    bl_metric_uu(a, r, th, guu);
    double om, ur;
    if (r < m.r_isco) {
        // the plunge inside the ISCO with the ISCO's energy and angular momentum
        const double wi = 1.0 / (std::pow(m.r_isco, 1.5) + a);
        double gi[4][4]; bl_metric_dd(a, m.r_isco, th, gi);
        double ui[4] = {1.0, 0.0, 0.0, wi};
        const double nrm = std::sqrt(-dot4(gi, ui, ui));
        for (double& x : ui) x /= nrm;
        double ud[4]; mv4(gi, ui, ud);
        const double e = ud[0], l = ud[3];
        const double Kc = guu[0][0] * e * e + 2.0 * guu[0][3] * e * l + guu[3][3] * l * l;
        const double urk = -std::sqrt(std::max(0.0, -(1.0 + Kc) / guu[1][1]));
        const double cov[4] = {e, urk, 0.0, l};
        double ut[4]; mv4(guu, cov, ut);
        om = ut[3] / ut[0];
        ur = ut[1];
    } else {
        om = 1.0 / (std::pow(r, 1.5) + a);
        ur = 0.0;
    }
    const double K = gdd[0][0] + 2.0 * om * gdd[0][3] + om * om * gdd[3][3];
    const double ut = std::sqrt(std::max(0.0, -(1.0 + ur * ur * gdd[1][1]) / K));
    const double u[4] = {ut, ur, 0.0, om * ut};
    // toroidal field made orthogonal to u and normalised to bmag
    const double Bc[4] = {0.0, 0.0, 0.0, 1.0};
    const double BdU = dot4(gdd, Bc, u);
    double b[4];
    for (int i = 0; i < 4; ++i) b[i] = Bc[i] + BdU * u[i];
    const double bn = std::sqrt(dot4(gdd, b, b));
    for (double& x : b) x *= bmag / bn;
    // ZAMO components of u (the spatial part is the fluid's ZAMO 4-velocity) and the field in the fluid frame
    const Zamo<double> z = zamo(a, r, th);
    double uz[4], bz[4], bf[4];
    bl_to_zamo(z, u, uz);
    bl_to_zamo(z, b, bz);
    const double ut3[3] = {uz[1], uz[2], uz[3]};
    double Lb[4][4]; boost(ut3, Lb);
    mat4vec(Lb, bz, bf);
    const double Bf[3] = {bf[1], bf[2], bf[3]};
    fr = local_frame(a, r, th, rc.eta, rc.lam, s.nur, s.nuth, rc.alpha, rc.beta, rc.thetao, ut3, Bf);
    const double nuf = nu_obs / fr.g;
    const double thB = std::acos(std::min(std::max(fr.cosB, -1.0), 1.0));
    cf = thermal_synchrotron(n, Th, bmag, nuf, thB, false, m.pandya);
    return true;
}

// ---- shells with given fluid-frame coefficients (test_polarized.jl:33-51) -----------------------------------
struct Slab { double r1, r2; Coef<double> c; double u[3], B[3]; };
inline bool slab_element(const Slab& e, const RayCtx& rc, const Sample<double>& s, Coef<double>& cf, Frame<double>& fr) {
    if (!(e.r1 <= s.r && s.r <= e.r2)) return false;
    fr = local_frame(rc.a, s.r, s.th, rc.eta, rc.lam, s.nur, s.nuth, rc.alpha, rc.beta, rc.thetao, e.u, e.B);
    cf = e.c;
    return true;
}

// ---- Gold et al. (2020) torus (test_gold2020.jl:14-53) -------------------------------------------------------
struct Gold { double A, alpha, h, l0, nup, rhou; };
// the five models: spin and (A, alpha, h, l0); model 2 runs at a = 1e-3 (a = 0 is outside the closed forms)
inline void gold_model(int i, double& a, Gold& g) {
    static const double P[5][5] = {{0.9, 0.0, -3.0, 0.0, 0.0}, {1e-3, 0.0, -2.0, 0.0, 1.0}, {0.9, 0.0, 0.0, 10.0 / 3.0, 1.0},
                                   {0.9, 1e5, 0.0, 10.0 / 3.0, 1.0}, {0.9, 1e6, 0.0, 100.0 / 3.0, 1.0}};
    a = P[i][0];
    g = {P[i][1], P[i][2], P[i][3], P[i][4], 230e9, 3e-18};
}
// fluid-frame j, alpha and the redshift at one sample (test_gold2020.jl:39-53)
inline void gold_coefficients(const Gold& m, const RayCtx& rc, const Sample<double>& s, double nu_obs, double& j, double& al, double& g) {
    const double r = s.r, th = s.th;
    const double nexp = (r / 10.0) * (r / 10.0) / 2.0 + (m.h * std::cos(th)) * (m.h * std::cos(th)) / 2.0;
    const double n = nexp < 200.0 ? m.rhou * std::exp(-nexp) : 0.0;
    const double R = r * std::sin(th);
    const double l = m.l0 / (1.0 + R) * std::pow(R, 1.5);
    double guu[4][4]; bl_metric_uu(rc.a, r, th, guu);
    const double ubar = std::sqrt(-1.0 / (guu[0][0] - 2.0 * guu[0][3] * l + guu[3][3] * l * l));
    const double ut = ubar * (-guu[0][0] + guu[0][3] * l);
    const double up = ubar * (-guu[0][3] + guu[3][3] * l);
    g = 1.0 / (ut - rc.lam * up);
    const double nuf = nu_obs / g;
    j = n * std::pow(nuf / m.nup, -m.alpha);
    al = m.A * n * std::pow(nuf / m.nup, -(2.5 + m.alpha)) + 1e-54;
}

}  // namespace kso
