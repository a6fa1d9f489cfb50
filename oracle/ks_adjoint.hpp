// The dual sweep: the gradient of a linear functional of the polarized image with respect to the
// thermal splat parameters, for the CPU oracle (test infrastructure; see ks_scalar.hpp).
//
// Restates src/Splats/dual_gradient.jl:1-249 and sample_adjoint (transport.jl:332-341). Along a ray
// the invariant Stokes vector arriving at the observer is T_1 with T_{N+1} = 0,
// T_k = E_k + O_k T_{k+1}; for l = w_1 . T_1 the adjoints of sample k are dl/dO_k = w_k T_{k+1}^T,
// dl/dE_k = w_k with w_{k+1} = O_k^T w_k. Per sample: j_bar = Ej^T w exactly; alpha_bar, rho_bar by
// 7-partial forward duals through the step operator; each splat's share through 7-partial duals in
// its fluid rows (logTe, logB, thB, phB, u1, u2, u3) and, since every coefficient is proportional
// to the density, through (c_bar . c) times the gradient of ln G (13 partials: rows 1-12 and the
// pattern rate) for the geometric rows and ln n_e (dual_gradient.jl:46-76).
#pragma once
#include "ks_splats.hpp"
#include "ks_dual.hpp"
#include <vector>

namespace kso {

using D7 = Dual<7>;
using D13 = Dual<13>;

// sample_adjoint (transport.jl:332-341)
inline void sample_adjoint(const double j4[4], const double a4[4], const double r3[3], double D, const double w[4], const double R[4],
                           M4<double>& O, double E[4], double jbar[4], double abar[4], double rbar[3]) {
    D7 ad[4], rd[3];
    for (int i = 0; i < 4; ++i) ad[i] = D7::seed(a4[i], i);
    for (int i = 0; i < 3; ++i) rd[i] = D7::seed(r3[i], 4 + i);
    M4<D7> Od, Ejd;
    step_operator<D7>(ad, rd, D7(D), Od, Ejd);
    D7 q(0.0);
    for (int i = 0; i < 4; ++i) {
        D7 oR(0.0), ej(0.0);
        for (int k = 0; k < 4; ++k) { oR = oR + Od.m[i][k] * R[k]; ej = ej + Ejd.m[i][k] * j4[k]; }
        q = q + w[i] * (oR + ej);
    }
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) O.m[i][k] = Od.m[i][k].v;
    for (int i = 0; i < 4; ++i) {
        double e = 0.0;
        for (int k = 0; k < 4; ++k) e += Ejd.m[i][k].v * j4[k];
        E[i] = e;
    }
    for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += Ejd.m[i][k].v * w[i];
        jbar[k] = s;
    }
    for (int i = 0; i < 4; ++i) abar[i] = q.d[i];
    for (int i = 0; i < 3; ++i) rbar[i] = q.d[4 + i];
}

// one thermal splat's share of sample s (dual_gradient.jl:46-76); grad is the splat's 21-row column
inline void element_adjoint(const Model& m, int i, const RayCtx& rc, const Sample<double>& s, double t_obs, double nu,
                            const double jbar[4], const double abar[4], const double rbar[3], double* grad) {
    double x, y, z;
    quasi_cartesian_ks(rc.a, s.r, s.th, s.ph, x, y, z);
    double t = t_obs - s.t;
    if (outside_support(m, i, t, x, y, z)) return;
    // weight with 13 partials: rows 0..11 and the pattern rate (last row)
    const int nr = m.nrows;
    std::vector<double> colv(nr);
    for (int r = 0; r < nr; ++r) colv[r] = m.P(r, i);
    D13 pr[32];
    for (int r = 0; r < nr; ++r) {
        int k = r <= 11 ? r : (r == nr - 1 ? 12 : -1);
        pr[r] = k >= 0 ? D13::seed(colv[r], k) : D13(colv[r]);
    }
    // pattern offset and Gaussian in duals (Splats.jl:58-93)
    D13 ph = pr[nr - 1] * (D13(t) - pr[10]);
    D13 sp = ksin(ph), cp = kcos(ph);
    D13 cx = pr[0] * cp - pr[1] * sp, cy = pr[0] * sp + pr[1] * cp;
    D13 d0 = D13(x) - cx, d1 = D13(y) - cy, d2 = D13(z) - pr[2];
    D13 e0 = d0 * cp + d1 * sp, e1 = -d0 * sp + d1 * cp, e2 = d2;
    D13 R[3][3];
    quat_rot(pr[6], pr[7], pr[8], pr[9], R);
    D13 u0 = R[0][0] * e0 + R[1][0] * e1 + R[2][0] * e2;
    D13 u1 = R[0][1] * e0 + R[1][1] * e1 + R[2][1] * e2;
    D13 u2 = R[0][2] * e0 + R[1][2] * e1 + R[2][2] * e2;
    D13 q0 = u0 * kexp(-pr[3]), q1 = u1 * kexp(-pr[4]), q2 = u2 * kexp(-pr[5]);
    D13 tt = (D13(t) - pr[10]) * kexp(-pr[11]);
    D13 Gd = kexp(-(q0 * q0 + q1 * q1 + q2 * q2 + tt * tt) / 2.0);
    double G = Gd.v;
    if (!(G > WEIGHT_CUTOFF)) return;
    double dlnG[13];
    for (int k = 0; k < 13; ++k) dlnG[k] = Gd.d[k] / G;
    // fluid rows with 7 partials (rows 13..19)
    D7 fl[7];
    for (int k = 0; k < 7; ++k) fl[k] = D7::seed(colv[13 + k], k);
    D7 Th = kexp(fl[0]);
    D7 Bmag = kexp(fl[1]);
    D7 st = ksin(fl[2]), ct = kcos(fl[2]), sph = ksin(fl[3]), cph = kcos(fl[3]);
    D7 B[3] = {Bmag * st * cph, Bmag * st * sph, Bmag * ct};
    D7 u[3] = {fl[4], fl[5], fl[6]};
    Frame<D7> fr = local_frame<D7>(D7(rc.a), D7(s.r), D7(s.th), D7(rc.eta), D7(rc.lam), s.nur, s.nuth, D7(rc.alpha), D7(rc.beta),
                                   D7(rc.thetao), u, B);
    D7 nuf = D7(nu) / fr.g;
    D7 thB = kacos(kmin(kmax(fr.cosB, D7(-1.0)), D7(1.0)));
    // Bessel K0, K1 at 1/Theta_e with their derivatives (K0' = -K1, K1' = -K0 - K1/x)
    const double* kb = m.bessel + 3 * i;
    D7 xb = D7(1.0) / Th;
    D7 k0 = chain(kb[0], -kb[1], xb), k1 = chain(kb[1], -kb[0] - kb[1] / xb.v, xb);
    D7 k2 = k0 + 2.0 * k1 / xb;
    double ne = std::exp(colv[12]) * G;
    Coef<D7> cf = thermal_synchrotron_k<D7>(D7(ne), Th, Bmag, nuf, thB, k0, k1, k2);
    Screen<D7> sc;
    if (!screen_coefficients<D7>(cf, fr, D7(nu), sc)) return;
    D7 q(0.0);
    for (int k = 0; k < 4; ++k) q = q + jbar[k] * sc.j[k] + abar[k] * sc.a[k];
    for (int k = 0; k < 3; ++k) q = q + rbar[k] * sc.r[k];
    double sv = q.v;
    for (int r = 0; r < 12; ++r) grad[r] += sv * dlnG[r];
    grad[12] += sv;
    for (int k = 0; k < 7; ++k) grad[13 + k] += q.d[k];
    grad[nr - 1] += sv * dlnG[12];
}

// Gradient of sum_pixels dstokes . observed_stokes over one ray (thermal splats); adds into grad
// (21 x ns, column-major) and returns the ray's observed Stokes vector.
inline void polarized_gradient_ray(const Model& m, double a, double thetao, double alpha, double beta, int N, int M, double t_obs, double nu,
                                   double L, int nmax, double slab, const double dstokes[4], double* grad, double image[4]) {
    Pixel<double> p = pixel_constants<double>(a, alpha, beta, thetao);
    RayCtx rc{a, alpha, beta, thetao, p.eta, p.lam};
    std::vector<Sample<double>> S(N);
    double dtau = 0, rt, rp;
    auto store = [&](int k, const Sample<double>& s, double dt) { S[k - 1] = s; dtau = dt; };
    march_ray(p, N, M, store, rt, rp);
    double hor = (1.0 + std::sqrt(1.0 - a * a)) * (1.0 + 1e-3);
    // half-orbit cutoff: first sample beyond the nmax-th passage (dual_gradient.jl:81-97)
    int kstop = N + 1;
    if (nmax >= 0) {
        Winding<double> w = winding_zero<double>();
        for (int k = 1; k <= N; ++k) {
            const Sample<double>& s = S[k - 1];
            if (!(s.ok && s.r > hor)) continue;
            wind(w, s.r * std::cos(s.th), slab);
            if (w.n > nmax) { kstop = k; break; }
        }
    }
    // per-sample steps, then tails back to front (dual_gradient.jl:129-147)
    struct Step { bool on; double j4[4], a4[4], r3[3], D; M4<double> O; double E[4]; };
    std::vector<Step> st(N);
    for (int k = 1; k <= N; ++k) {
        Step& q = st[k - 1];
        q.on = false;
        if (k >= kstop) continue;
        const Sample<double>& s = S[k - 1];
        if (!(s.ok && s.r > hor)) continue;
        if (!accumulate_elements(m, rc, s, t_obs, nu, q.j4, q.a4, q.r3)) continue;
        double c = std::cos(s.th);
        q.D = L / (nu / NU0) * (s.r * s.r + a * a * c * c) * dtau;
        M4<double> Ej;
        step_operator(q.a4, q.r3, q.D, q.O, Ej);
        for (int i = 0; i < 4; ++i) q.E[i] = Ej.m[i][0] * q.j4[0] + Ej.m[i][1] * q.j4[1] + Ej.m[i][2] * q.j4[2] + Ej.m[i][3] * q.j4[3];
        q.on = true;
    }
    std::vector<double> tails(4 * (size_t)(N + 1), 0.0);
    for (int k = N; k >= 1; --k) {
        double* Tk = &tails[4 * (k - 1)];
        const double* Tn = &tails[4 * k];
        const Step& q = st[k - 1];
        if (q.on) for (int i = 0; i < 4; ++i) Tk[i] = q.E[i] + q.O.m[i][0] * Tn[0] + q.O.m[i][1] * Tn[1] + q.O.m[i][2] * Tn[2] + q.O.m[i][3] * Tn[3];
        else for (int i = 0; i < 4; ++i) Tk[i] = Tn[i];
    }
    double nh3 = std::pow(nu / NU0, 3);
    for (int i = 0; i < 4; ++i) image[i] = tails[i] * nh3;
    // forward adjoint sweep (dual_gradient.jl:151-178)
    double w[4];
    for (int i = 0; i < 4; ++i) w[i] = dstokes[i] * nh3;
    for (int k = 1; k <= N; ++k) {
        const Step& q = st[k - 1];
        if (!q.on) continue;
        M4<double> O; double E[4], jb[4], ab[4], rb[3];
        sample_adjoint(q.j4, q.a4, q.r3, q.D, w, &tails[4 * k], O, E, jb, ab, rb);
        for (int i = 0; i < m.ns; ++i) element_adjoint(m, i, rc, S[k - 1], t_obs, nu, jb, ab, rb, grad + (size_t)m.nrows * i);
        double wn[4];
        for (int c = 0; c < 4; ++c) wn[c] = O.m[0][c] * w[0] + O.m[1][c] * w[1] + O.m[2][c] * w[2] + O.m[3][c] * w[3];
        for (int c = 0; c < 4; ++c) w[c] = wn[c];
    }
}

}  // namespace kso
