// The dual sweep: the gradient of a linear functional of the polarized image with respect to the
// thermal splat parameters, for the CPU oracle (test infrastructure; see ks_scalar.hpp).
//
// Restates src/Splats/dual_gradient.jl:1-249 and sample_adjoint (transport.jl:198-207). Along a ray
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

// sample_adjoint (transport.jl:198-207); S = double, or Counted to measure the work. The 7 partials of
// q = w . (O R + Ej j) with respect to (alpha, rho) come from duals through the step's polynomial
// coefficients, contracted as sum_n o_n w.(K'^n R) + Dp e_n w.(K'^n j) (the same function as the
// ForwardDiff derivative of the assembled matrices, without forming 4x4 dual matrices). Returns
// O^T w (the next adjoint), E = Ej j, j_bar = Ej^T w, alpha_bar and rho_bar.
template <class S, bool SPLIT = false>
inline void sample_adjoint(const S j4[4], const S a4[4], const S r3[3], S D, const S w[4], const S R[4],
                           S Otw[4], S E[4], S jbar[4], S abar[4], S rbar[3]) {
    if constexpr (SPLIT) {
        // the device's form: four-partial duals over the invariants (alpha_I, a2, r2, ar) of the scaled K
        // through the coefficients, six-partial duals over the scaled K' in the products w.(K'^n v), joined
        // by the chain rule; the bars are the derivatives with respect to the scaled entries times Dp
        using D4 = Dual<4, S>;
        using D6 = Dual<6, S>;
        S Dp = pow2_below(D), f = D / Dp;
        S K6v[6];
        S aIs = a4[0] * Dp;
        for (int i = 0; i < 3; ++i) { K6v[i] = a4[1 + i] * Dp; K6v[3 + i] = r3[i] * Dp; }
        S a2 = K6v[0] * K6v[0] + K6v[1] * K6v[1] + K6v[2] * K6v[2];
        S r2 = K6v[3] * K6v[3] + K6v[4] * K6v[4] + K6v[5] * K6v[5];
        S ar = K6v[0] * K6v[3] + K6v[1] * K6v[4] + K6v[2] * K6v[5];
        D4 o[4], e4[4];
        step_coefficients_inv<D4>(D4::seed(aIs, 0), D4::seed(a2, 1), D4::seed(r2, 2), D4::seed(ar, 3), D4(f), o, e4);
        D6 K6[6];
        for (int i = 0; i < 6; ++i) K6[i] = D6::seed(K6v[i], i);
        D6 vR[4], vj[4], t4[4], sR[4], sj[4];
        S sR0(0.0), sj0(0.0);
        for (int i = 0; i < 4; ++i) { sR0 = sR0 + w[i] * R[i]; sj0 = sj0 + w[i] * j4[i]; vR[i] = D6(R[i]); vj[i] = D6(j4[i]); }
        sR[0] = D6(sR0); sj[0] = D6(sj0);
        for (int n = 1; n < 4; ++n) {
            kprime_mul(K6, vR, t4); for (int i = 0; i < 4; ++i) vR[i] = t4[i];
            kprime_mul(K6, vj, t4); for (int i = 0; i < 4; ++i) vj[i] = t4[i];
            sR[n] = vR[0] * D6(w[0]) + vR[1] * D6(w[1]) + vR[2] * D6(w[2]) + vR[3] * D6(w[3]);
            sj[n] = vj[0] * D6(w[0]) + vj[1] * D6(w[1]) + vj[2] * D6(w[2]) + vj[3] * D6(w[3]);
        }
        S gi[4], hm[6];
        for (int k = 0; k < 4; ++k) {
            S acc(0.0);
            for (int n = 0; n < 4; ++n) acc = acc + o[n].d[k] * sR[n].v + Dp * (e4[n].d[k] * sj[n].v);
            gi[k] = acc;
        }
        for (int m = 0; m < 6; ++m) {
            S acc(0.0);
            for (int n = 1; n < 4; ++n) acc = acc + o[n].v * sR[n].d[m] + Dp * (e4[n].v * sj[n].d[m]);
            hm[m] = acc;
        }
        abar[0] = gi[0] * Dp;
        for (int i = 0; i < 3; ++i) {
            abar[1 + i] = (hm[i] + 2.0 * K6v[i] * gi[1] + K6v[3 + i] * gi[3]) * Dp;
            rbar[i] = (hm[3 + i] + 2.0 * K6v[3 + i] * gi[2] + K6v[i] * gi[3]) * Dp;
        }
        S u[4], y[4], z[4], v[4];
        for (int i = 0; i < 4; ++i) {
            u[i] = w[i]; v[i] = j4[i];
            Otw[i] = o[0].v * w[i]; jbar[i] = Dp * e4[0].v * w[i]; E[i] = Dp * e4[0].v * j4[i];
        }
        for (int n = 1; n < 4; ++n) {
            kprime_tmul(K6v, u, y); kprime_mul(K6v, v, z);
            for (int i = 0; i < 4; ++i) {
                u[i] = y[i]; v[i] = z[i];
                Otw[i] = Otw[i] + o[n].v * u[i]; jbar[i] = jbar[i] + Dp * e4[n].v * u[i]; E[i] = E[i] + Dp * e4[n].v * v[i];
            }
        }
        return;
    }
    using D7 = Dual<7, S>;
    D7 ad[4], rd[3];
    for (int i = 0; i < 4; ++i) ad[i] = D7::seed(a4[i], i);
    for (int i = 0; i < 3; ++i) rd[i] = D7::seed(r3[i], 4 + i);
    D7 Dp, K6[6], o[4], e4[4];
    step_coefficients<D7>(ad, rd, D7(D), Dp, K6, o, e4);
    S wR(0.0), wj(0.0);
    for (int i = 0; i < 4; ++i) { wR = wR + w[i] * R[i]; wj = wj + w[i] * j4[i]; }
    D7 q = o[0] * D7(wR) + Dp * e4[0] * D7(wj);
    D7 vR[4], vj[4], t4[4];
    for (int i = 0; i < 4; ++i) { vR[i] = D7(R[i]); vj[i] = D7(j4[i]); }
    for (int n = 1; n < 4; ++n) {
        kprime_mul(K6, vR, t4); for (int i = 0; i < 4; ++i) vR[i] = t4[i];
        kprime_mul(K6, vj, t4); for (int i = 0; i < 4; ++i) vj[i] = t4[i];
        D7 sR = vR[0] * D7(w[0]) + vR[1] * D7(w[1]) + vR[2] * D7(w[2]) + vR[3] * D7(w[3]);
        D7 sj = vj[0] * D7(w[0]) + vj[1] * D7(w[1]) + vj[2] * D7(w[2]) + vj[3] * D7(w[3]);
        q = q + o[n] * sR + Dp * e4[n] * sj;
    }
    for (int i = 0; i < 4; ++i) abar[i] = q.d[i];
    for (int i = 0; i < 3; ++i) rbar[i] = q.d[4 + i];
    // values: O^T w, Ej^T w and E = Ej j by repeated products with K'
    S K6v[6], u[4], y[4], z[4], v[4];
    for (int i = 0; i < 6; ++i) K6v[i] = K6[i].v;
    const S Dpv = Dp.v;
    for (int i = 0; i < 4; ++i) {
        u[i] = w[i]; v[i] = j4[i];
        Otw[i] = o[0].v * w[i]; jbar[i] = Dpv * e4[0].v * w[i]; E[i] = Dpv * e4[0].v * j4[i];
    }
    for (int n = 1; n < 4; ++n) {
        kprime_tmul(K6v, u, y); kprime_mul(K6v, v, z);
        for (int i = 0; i < 4; ++i) {
            u[i] = y[i]; v[i] = z[i];
            Otw[i] = Otw[i] + o[n].v * u[i]; jbar[i] = jbar[i] + Dpv * e4[n].v * u[i]; E[i] = E[i] + Dpv * e4[n].v * v[i];
        }
    }
}

// d ln G of the splat weight (Splats.jl:58-93) with respect to rows 1-12 and the pattern rate written
// out by hand: ln G = -(|q|^2 + tt^2)/2 with q_b = u_b e^{-s_b}, u = R(quaternion)^T e,
// e = Rz(-ph) (X - c(ph)), c(ph) = Rz(ph) (c0, c1) + c2 e_z, ph = omega (t - t0), tt = (t - t0) e^{-s_t}.
// The same function as the 13-partial duals (checked against them in tests/test_oracle_adjoint.py);
// it is how the GPU evaluates it, and the work model counts this form. Returns G (0 when culled).
template <class S>
inline S weight_dlog_analytic(const Model& m, int i, S t, S x, S y, S z, S dl[13]) {
    if (outside_support(m, i, t, x, y, z)) return S(0.0);
    const int nr = m.nrows;
    const double c0 = m.P(0, i), c1 = m.P(1, i), c2 = m.P(2, i), t0 = m.P(10, i), om = m.P(nr - 1, i);
    S ph = om * (t - t0);
    S sp = ksin(ph), cp = kcos(ph);
    S cx = c0 * cp - c1 * sp, cy = c0 * sp + c1 * cp;
    S d0 = x - cx, d1 = y - cy, d2 = z - c2;
    S e[3] = {d0 * cp + d1 * sp, -d0 * sp + d1 * cp, d2};
    S R[3][3];
    quat_rot(S(m.P(6, i)), S(m.P(7, i)), S(m.P(8, i)), S(m.P(9, i)), R);
    S q[3], is[3];
    for (int b = 0; b < 3; ++b) {
        is[b] = kexp(S(-m.P(3 + b, i)));
        q[b] = (R[0][b] * e[0] + R[1][b] * e[1] + R[2][b] * e[2]) * is[b];
    }
    S iw = kexp(S(-m.P(11, i)));
    S dt = t - t0, tt = dt * iw;
    S G = kexp(-(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + tt * tt) / 2.0);
    if (!(G > WEIGHT_CUTOFF)) return S(0.0);
    S gu[3], ge[3];
    for (int b = 0; b < 3; ++b) { gu[b] = -q[b] * is[b]; dl[3 + b] = q[b] * q[b]; }
    for (int a = 0; a < 3; ++a) ge[a] = R[a][0] * gu[0] + R[a][1] * gu[1] + R[a][2] * gu[2];
    // rotation: d ln G / dR[a][b] = e_a gu_b, through the normalised quaternion (w, x, y, z)
    S GR[9];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) GR[3 * a + b] = e[a] * gu[b];
    double qq[4] = {m.P(6, i), m.P(7, i), m.P(8, i), m.P(9, i)};
    double nq = std::sqrt(qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3]);
    S qw = S(qq[0] / nq), qx = S(qq[1] / nq), qy = S(qq[2] / nq), qz = S(qq[3] / nq);
    S gn[4];
    gn[0] = 2.0 * (-qz * GR[1] + qy * GR[2] + qz * GR[3] - qx * GR[5] - qy * GR[6] + qx * GR[7]);
    gn[1] = 2.0 * (qy * GR[1] + qz * GR[2] + qy * GR[3] - 2.0 * qx * GR[4] - qw * GR[5] + qz * GR[6] + qw * GR[7] - 2.0 * qx * GR[8]);
    gn[2] = 2.0 * (-2.0 * qy * GR[0] + qx * GR[1] + qw * GR[2] + qx * GR[3] + qz * GR[5] - qw * GR[6] + qz * GR[7] - 2.0 * qy * GR[8]);
    gn[3] = 2.0 * (-2.0 * qz * GR[0] - qw * GR[1] + qx * GR[2] + qw * GR[3] - 2.0 * qz * GR[4] + qy * GR[5] + qx * GR[6] + qy * GR[7]);
    S ng = qw * gn[0] + qx * gn[1] + qy * gn[2] + qz * gn[3];
    S qn[4] = {qw, qx, qy, qz};
    for (int k = 0; k < 4; ++k) dl[6 + k] = (gn[k] - qn[k] * ng) / nq;
    // centre: d ln G / dc = -Rz(ph) ge
    S gd0 = ge[0] * cp - ge[1] * sp, gd1 = ge[0] * sp + ge[1] * cp;
    dl[0] = -(gd0 * cp + gd1 * sp);
    dl[1] = -(gd1 * cp - gd0 * sp);
    dl[2] = -ge[2];
    // pattern phase: de/dph = (e1 + c1, -(e0 + c0), 0)
    S dph = ge[0] * (e[1] + c1) - ge[1] * (e[0] + c0);
    dl[10] = -om * dph + tt * iw;
    dl[11] = tt * tt;
    dl[12] = dph * dt;
    return G;
}

// The device's form of the frame partials: local_frame's (g, cos B, chi) by compact expressions (the boost
// applied to the ZAMO momentum and, inverted, to the unit polarisation direction h = f / |f|) with their
// partials over (theta_B, phi_B, u1, u2, u3): g and the wave vector k forward in three-partial duals over u,
// cos B = k . bh from those, chi by one reverse sweep of its own chain joined to the partials of k and of
// q = 1 / (gamma + 1). chi depends on h only, so hbar . h = 0 and fbar = hbar / |f|. Checked through the
// whole gradient against the seven-partial duals of the other path.
template <class S> struct FrameJ { S g, cosB, chi, dg[5], dcosB[5], dchi[5]; };

template <class S>
FrameJ<S> frame_jacobian(const RayCtx& rc, const Sample<S>& s, S thB, S phB, const S uv[3]) {
    using D3 = Dual<3, S>;
    const S a = S(rc.a);
    S pd[4], pu[4], pz[4];
    momentum_bl_d(a, s.r, s.th, S(rc.eta), S(rc.lam), s.nur, s.nuth, pd);
    metric_uu_mul(a, s.r, s.th, pd, pu);
    const Zamo<S> zm = zamo(a, s.r, s.th);
    bl_to_zamo(zm, pu, pz);
    const S st = ksin(s.th), ct = kcos(s.th);
    const S sthB = ksin(thB), cthB = kcos(thB), sphB = ksin(phB), cphB = kcos(phB);
    const S bh[3] = {sthB * cphB, sthB * sphB, cthB};
    const S bt[3] = {cthB * cphB, cthB * sphB, -sthB}, bp[3] = {-sthB * sphB, sthB * cphB, S(0.0)};
    const D3 u[3] = {D3::seed(uv[0], 0), D3::seed(uv[1], 1), D3::seed(uv[2], 2)};
    const D3 gam = ksqrt(D3(1.0) + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const D3 q = D3(1.0) / (gam + D3(1.0));
    const D3 ud = u[0] * D3(pz[1]) + u[1] * D3(pz[2]) + u[2] * D3(pz[3]);
    const D3 c1 = ud * q - D3(pz[0]);
    const D3 g = D3(1.0) / (gam * D3(pz[0]) - ud);
    D3 k[3];
    for (int i = 0; i < 3; ++i) k[i] = (D3(pz[1 + i]) + u[i] * c1) * g;
    const D3 cosB = k[0] * D3(bh[0]) + k[1] * D3(bh[1]) + k[2] * D3(bh[2]);
    FrameJ<S> o;
    o.g = g.v; o.cosB = cosB.v; o.chi = S(0.0);
    o.dg[0] = o.dg[1] = S(0.0);
    o.dcosB[0] = k[0].v * bt[0] + k[1].v * bt[1] + k[2].v * bt[2];
    o.dcosB[1] = k[0].v * bp[0] + k[1].v * bp[1] + k[2].v * bp[2];
    for (int i = 0; i < 3; ++i) { o.dg[2 + i] = g.d[i]; o.dcosB[2 + i] = cosB.d[i]; }
    for (int i = 0; i < 5; ++i) o.dchi[i] = S(0.0);
    const S f[3] = {k[1].v * bh[2] - k[2].v * bh[1], k[2].v * bh[0] - k[0].v * bh[2], k[0].v * bh[1] - k[1].v * bh[0]};
    const S fn = ksqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fn > 0.0)) return o;
    const S ifn = 1.0 / fn;
    const S h[3] = {f[0] * ifn, f[1] * ifn, f[2] * ifn};
    const S uf = u[0].v * h[0] + u[1].v * h[1] + u[2].v * h[2];
    const S ufq = uf * q.v;
    const S C1 = a * st * st, C2 = s.r * s.r + a * a, act = a * ct;
    const S mu = -(S(rc.alpha) + a * ksin(S(rc.thetao))), beta = S(rc.beta);
    const S ft = uf / zm.lapse, fr = (h[0] + u[0].v * ufq) / zm.sgrr, fth = -(h[2] + u[2].v * ufq) / zm.sgthth;
    const S fph = (h[1] + u[1].v * ufq) / zm.sgpp + zm.omega * ft;
    const S A = pu[0] * fr - pu[1] * ft + C1 * (pu[1] * fph - pu[3] * fr);
    const S Bw = (C2 * (pu[3] * fth - pu[2] * fph) - a * (pu[0] * fth - pu[2] * ft)) * st;
    const S kk1 = A * s.r - Bw * act, kk2 = -(A * act + Bw * s.r);
    const S ea = beta * kk2 - mu * kk1, eb = beta * kk1 + mu * kk2;
    o.chi = katan2(-ea, eb);
    const S den = ea * ea + eb * eb;
    if (!(den > 0.0)) return o;
    const S eab = -eb / den, ebb = ea / den;
    const S kk1b = beta * ebb - mu * eab, kk2b = beta * eab + mu * ebb;
    const S Ab = s.r * kk1b - act * kk2b, Bws = -(act * kk1b + s.r * kk2b) * st;
    const S fthb = Bws * (C2 * pu[3] - a * pu[0]);
    const S fphb = Ab * (C1 * pu[1]) - Bws * (C2 * pu[2]);
    const S ftb = Bws * (a * pu[2]) - Ab * pu[1] + fphb * zm.omega;
    const S zb[3] = {(Ab * (pu[0] - C1 * pu[3])) / zm.sgrr, fphb / zm.sgpp, -fthb / zm.sgthth};
    const S ufqb = zb[0] * u[0].v + zb[1] * u[1].v + zb[2] * u[2].v;
    const S ufb = ftb / zm.lapse + ufqb * q.v;
    const S qb = ufqb * uf;
    S fb[3], ub[3];
    for (int i = 0; i < 3; ++i) { fb[i] = (zb[i] + ufb * u[i].v) * ifn; ub[i] = zb[i] * ufq + ufb * h[i]; }
    // f = k x bh: kbar = bh x fbar, bhbar = fbar x k
    const S kb[3] = {bh[1] * fb[2] - bh[2] * fb[1], bh[2] * fb[0] - bh[0] * fb[2], bh[0] * fb[1] - bh[1] * fb[0]};
    const S bb[3] = {fb[1] * k[2].v - fb[2] * k[1].v, fb[2] * k[0].v - fb[0] * k[2].v, fb[0] * k[1].v - fb[1] * k[0].v};
    o.dchi[0] = bb[0] * bt[0] + bb[1] * bt[1] + bb[2] * bt[2];
    o.dchi[1] = bb[0] * bp[0] + bb[1] * bp[1] + bb[2] * bp[2];
    for (int i = 0; i < 3; ++i) o.dchi[2 + i] = ub[i] + qb * q.d[i] + kb[0] * k[0].d[i] + kb[1] * k[1].d[i] + kb[2] * k[2].d[i];
    return o;
}

// one thermal splat's share of sample s (dual_gradient.jl:46-76); grad is the splat's 21-row column.
// ANALYTIC selects the device's form: the hand-written d ln G instead of the 13-partial duals, and the
// seven fluid partials in two passes joined by the chain rule instead of one 7-partial pass (the sweep
// passes it on to sample_adjoint, whose device form splits the same way).
template <class S, bool ANALYTIC = false>
inline void element_adjoint(const Model& m, int i, const RayCtx& rc, const Sample<S>& s, S t_obs, S nu,
                            const S jbar[4], const S abar[4], const S rbar[3], S* grad) {
    using D7 = Dual<7, S>;
    using D13 = Dual<13, S>;
    S x, y, z;
    quasi_cartesian_ks(S(rc.a), s.r, s.th, s.ph, x, y, z);
    S t = t_obs - s.t;
    const int nr = m.nrows;
    std::vector<double> colv(nr);
    for (int r = 0; r < nr; ++r) colv[r] = m.P(r, i);
    S G, dlnG[13];
    if constexpr (ANALYTIC) {
        G = weight_dlog_analytic(m, i, t, x, y, z, dlnG);
        if (!(G > 0.0)) return;
    } else {
    if (outside_support(m, i, t, x, y, z)) return;
    // weight with 13 partials: rows 0..11 and the pattern rate (last row)
    D13 pr[32];
    for (int r = 0; r < nr; ++r) {
        int k = r <= 11 ? r : (r == nr - 1 ? 12 : -1);
        pr[r] = k >= 0 ? D13::seed(S(colv[r]), k) : D13(colv[r]);
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
    G = Gd.v;
    if (!(G > WEIGHT_CUTOFF)) return;
    for (int k = 0; k < 13; ++k) dlnG[k] = Gd.d[k] / G;
    }
    const double* kb = m.bessel + 3 * i;
    S ne = kexp(S(colv[12])) * G;
    if constexpr (ANALYTIC) {
        // the device's form: the frame (g, cos B, chi) with five partials (theta_B, phi_B, u1..u3), the
        // contribution with four (ln Theta_e, ln B, g, cos B) and its chi derivative written out, joined by
        // the chain rule
        const S uv[3] = {S(colv[17]), S(colv[18]), S(colv[19])};
        const FrameJ<S> fr = frame_jacobian(rc, s, S(colv[15]), S(colv[16]), uv);
        using D4 = Dual<4, S>;
        D4 Th = kexp(D4::seed(S(colv[13]), 0));
        D4 Bmag = kexp(D4::seed(S(colv[14]), 1));
        D4 g = D4::seed(fr.g, 2), cosB = D4::seed(fr.cosB, 3);
        D4 nuf = D4(nu) / g;
        D4 thB = kacos(kmin(kmax(cosB, D4(-1.0)), D4(1.0)));
        D4 xb = D4(1.0) / Th;
        D4 k0 = chain(S(kb[0]), S(-kb[1]), xb), k1 = chain(S(kb[1]), S(-kb[0] - kb[1] / xb.v), xb);
        D4 k2 = k0 + 2.0 * k1 / xb;
        Coef<D4> cf = thermal_synchrotron_k<D4>(D4(ne), Th, Bmag, nuf, thB, k0, k1, k2);
        if (!(cf.jI > 0.0 || cf.aI > 0.0 || val(cf.rQ) != 0.0 || val(cf.rV) != 0.0)) return;
        // the screen coefficients before the rotation by 2 chi, whose derivative is written out
        Coef<D4> ci = invariants(cap_polarization(cf), nuf);
        D4 A = D4(jbar[0]) * ci.jI + D4(jbar[3]) * ci.jV + D4(abar[0]) * ci.aI + D4(abar[3]) * ci.aV + D4(rbar[2]) * ci.rV;
        D4 Bc = D4(jbar[1]) * ci.jQ + D4(abar[1]) * ci.aQ + D4(rbar[0]) * ci.rQ;
        D4 Bs = D4(jbar[2]) * ci.jQ + D4(abar[2]) * ci.aQ + D4(rbar[1]) * ci.rQ;
        S s2 = ksin(2.0 * fr.chi), c2 = kcos(2.0 * fr.chi);
        D4 q = A + Bc * D4(c2) + Bs * D4(s2);
        S qchi = 2.0 * (Bs.v * c2 - Bc.v * s2);
        S sv = q.v;
        for (int r = 0; r < 12; ++r) grad[r] = grad[r] + sv * dlnG[r];
        grad[12] = grad[12] + sv;
        grad[13] = grad[13] + q.d[0];
        grad[14] = grad[14] + q.d[1];
        for (int k = 0; k < 5; ++k) grad[15 + k] = grad[15 + k] + q.d[2] * fr.dg[k] + q.d[3] * fr.dcosB[k] + qchi * fr.dchi[k];
        grad[nr - 1] = grad[nr - 1] + sv * dlnG[12];
    } else {
        // fluid rows with 7 partials (rows 13..19)
        D7 fl[7];
        for (int k = 0; k < 7; ++k) fl[k] = D7::seed(S(colv[13 + k]), k);
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
        D7 xb = D7(1.0) / Th;
        D7 k0 = chain(S(kb[0]), S(-kb[1]), xb), k1 = chain(S(kb[1]), S(-kb[0] - kb[1] / xb.v), xb);
        D7 k2 = k0 + 2.0 * k1 / xb;
        Coef<D7> cf = thermal_synchrotron_k<D7>(D7(ne), Th, Bmag, nuf, thB, k0, k1, k2);
        Screen<D7> sc;
        if (!screen_coefficients<D7>(cf, fr, D7(nu), sc)) return;
        D7 q(0.0);
        for (int k = 0; k < 4; ++k) q = q + D7(jbar[k]) * sc.j[k] + D7(abar[k]) * sc.a[k];
        for (int k = 0; k < 3; ++k) q = q + D7(rbar[k]) * sc.r[k];
        S sv = q.v;
        for (int r = 0; r < 12; ++r) grad[r] = grad[r] + sv * dlnG[r];
        grad[12] = grad[12] + sv;
        for (int k = 0; k < 7; ++k) grad[13 + k] = grad[13 + k] + q.d[k];
        grad[nr - 1] = grad[nr - 1] + sv * dlnG[12];
    }
}

// One element's share of one sample for any fluid population, with or without knots: every row of the splat
// seeded as a dual number (up to 22 partials) through the same element_rows and screen coefficients that the
// transport evaluates, q = j_bar . j + alpha_bar . alpha + rho_bar . rho; adds dq/d(row) into grad. It restates
// nothing by hand (no chain rule written out), so it checks the device's hand-split element adjoints. The 2F1
// factor of kappa splats is differentiated with the splat's kappa and w (recomputed in duals), not held.
inline void element_adjoint_generic(const Model& m, int i, const RayCtx& rc, const Sample<double>& s, double t_obs, double nu,
                                    const double jbar[4], const double abar[4], const double rbar[3], double* grad) {
    using DG = Dual<22, double>;
    DG pr[32];
    for (int r = 0; r < m.nrows; ++r) pr[r] = DG::seed(m.P(r, i), r);
    const Sample<DG> sd{DG(s.t), DG(s.r), DG(s.th), DG(s.ph), s.nur, s.nuth, s.ok};
    Coef<DG> cf; Frame<DG> fr;
    if (!element_rows(m, i, rc, sd, DG(t_obs), DG(nu), pr, cf, fr)) return;
    Screen<DG> sc;
    if (!screen_coefficients<DG>(cf, fr, DG(nu), sc)) return;
    DG q(0.0);
    for (int k = 0; k < 4; ++k) q = q + DG(jbar[k]) * sc.j[k] + DG(abar[k]) * sc.a[k];
    for (int k = 0; k < 3; ++k) q = q + DG(rbar[k]) * sc.r[k];
    for (int r = 0; r < m.nrows; ++r) grad[r] = grad[r] + q.d[r];
}

// Gradient of sum_pixels dstokes . observed_stokes over one ray; adds into grad (21 x ns, column-major, thermal
// splats; with GENERIC any population, knots and composites, nrows x ns per part in order) and returns the ray's observed Stokes vector. With S = Counted, phases (if
// given) receives the operation counts at the ends of the march, the forward steps, the tails and the
// adjoint sweep, and phases[4] the part of the sweep spent in sample_adjoint.
template <class S, bool ANALYTIC = false, bool GENERIC = false>
inline void polarized_gradient_ray(const Model& m, double a, double thetao, double alpha, double beta, int N, int M, double t_obs, double nu,
                                   double L, int nmax, double slab, const double dstokes[4], S* grad, S image[4], OpCounts* phases = nullptr) {
    Pixel<S> p = pixel_constants<S>(S(a), S(alpha), S(beta), S(thetao));
    RayCtx rc{a, alpha, beta, thetao, val(p.eta), val(p.lam)};
    std::vector<Sample<S>> Sm(N);
    S dtau(0.0), rt, rp;
    auto store = [&](int k, const Sample<S>& s, S dt) { Sm[k - 1] = s; dtau = dt; };
    march_ray(p, N, M, store, rt, rp);
    if (phases) phases[0] = op_counts();
    double hor = (1.0 + std::sqrt(1.0 - a * a)) * (1.0 + 1e-3);
    // half-orbit cutoff: first sample beyond the nmax-th passage (dual_gradient.jl:81-97)
    int kstop = N + 1;
    if (nmax >= 0) {
        Winding<S> w = winding_zero<S>();
        for (int k = 1; k <= N; ++k) {
            const Sample<S>& s = Sm[k - 1];
            if (!(s.ok && s.r > hor)) continue;
            wind(w, s.r * kcos(s.th), S(slab));
            if (w.n > nmax) { kstop = k; break; }
        }
    }
    // per-sample steps, then tails back to front (dual_gradient.jl:129-147)
    struct Step { bool on; S j4[4], a4[4], r3[3], D; M4<S> O; S E[4]; };
    std::vector<Step> st(N);
    const S tS(t_obs), nuS(nu);
    for (int k = 1; k <= N; ++k) {
        Step& q = st[k - 1];
        q.on = false;
        if (k >= kstop) continue;
        const Sample<S>& s = Sm[k - 1];
        if (!(s.ok && s.r > hor)) continue;
        if (!accumulate_elements(m, rc, s, tS, nuS, q.j4, q.a4, q.r3)) continue;
        S c = kcos(s.th);
        q.D = L / (nu / NU0) * (s.r * s.r + a * a * c * c) * dtau;
        M4<S> Ej;
        step_operator(q.a4, q.r3, q.D, q.O, Ej);
        for (int i = 0; i < 4; ++i) q.E[i] = Ej.m[i][0] * q.j4[0] + Ej.m[i][1] * q.j4[1] + Ej.m[i][2] * q.j4[2] + Ej.m[i][3] * q.j4[3];
        q.on = true;
    }
    if (phases) phases[1] = op_counts();
    std::vector<S> tails(4 * (size_t)(N + 1), S(0.0));
    for (int k = N; k >= 1; --k) {
        S* Tk = &tails[4 * (k - 1)];
        const S* Tn = &tails[4 * k];
        const Step& q = st[k - 1];
        if (q.on) for (int i = 0; i < 4; ++i) Tk[i] = q.E[i] + q.O.m[i][0] * Tn[0] + q.O.m[i][1] * Tn[1] + q.O.m[i][2] * Tn[2] + q.O.m[i][3] * Tn[3];
        else for (int i = 0; i < 4; ++i) Tk[i] = Tn[i];
    }
    double nh3 = std::pow(nu / NU0, 3);
    for (int i = 0; i < 4; ++i) image[i] = tails[i] * nh3;
    if (phases) phases[2] = op_counts();
    // forward adjoint sweep (dual_gradient.jl:151-178)
    S w[4];
    for (int i = 0; i < 4; ++i) w[i] = S(dstokes[i] * nh3);
    OpCounts in_sample;
    for (int k = 1; k <= N; ++k) {
        const Step& q = st[k - 1];
        if (!q.on) continue;
        S Otw[4], E[4], jb[4], ab[4], rb[3];
        const OpCounts before = op_counts();
        sample_adjoint<S, ANALYTIC>(q.j4, q.a4, q.r3, q.D, w, &tails[4 * k], Otw, E, jb, ab, rb);
        in_sample.flops += op_counts().flops - before.flops;
        in_sample.special += op_counts().special - before.special;
        if constexpr (GENERIC) {
            // every part of a composite, its gradient block after the previous parts' (nrows x ns each)
            size_t off = 0;
            for (const Model* pm = &m; pm; pm = pm->next) {
                for (int i = 0; i < pm->ns; ++i)
                    element_adjoint_generic(*pm, i, rc, Sm[k - 1], t_obs, nu, jb, ab, rb, grad + off + (size_t)pm->nrows * i);
                off += (size_t)pm->nrows * pm->ns;
            }
        } else {
            for (int i = 0; i < m.ns; ++i) element_adjoint<S, ANALYTIC>(m, i, rc, Sm[k - 1], tS, nuS, jb, ab, rb, grad + (size_t)m.nrows * i);
        }
        for (int c = 0; c < 4; ++c) w[c] = Otw[c];
    }
    if (phases) { phases[3] = op_counts(); phases[4] = in_sample; }
}

}  // namespace kso
