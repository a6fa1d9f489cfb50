// Device pieces of the dual-sweep adjoint (thermal splats).
//
// Reference: sample_adjoint (src/Transfer/transport.jl:198-207) and element_adjoint! / ParamColumn
// (src/Splats/dual_gradient.jl:22-76). Along a ray, l = w_1 . T_1 with T_k = E_k + O_k T_{k+1}; the
// adjoints of sample k are dl/dO_k = w_k T_{k+1}^T and dl/dE_k = w_k, w_{k+1} = O_k^T w_k.
//   * sample_adjoint: j_bar = Ej^T w exactly; alpha_bar, rho_bar through the step's polynomial
//     coefficients (O = sum o_n K'^n, Ej = Dp sum e_n K'^n), contracted as sum_n o_n w.(K'^n T) +
//     Dp e_n w.(K'^n j) so that no 4x4 dual matrix is formed: four-partial duals over the invariants the
//     coefficients depend on, six-partial duals over K' in the products, joined by the chain rule.
//   * element_adjoint_thermal: one splat's share of one sample. Every coefficient is proportional to
//     the density n_e G, so the 13 geometric rows (centre, widths, orientation, t0, temporal width,
//     pattern rate) and ln n_e enter through (c_bar . c) d ln G, with d ln G written out by hand
//     below; the 7 fluid rows (ln Theta_e, ln B, field angles, u) through 7-partial duals of the
//     frame and the thermal coefficients.
#pragma once
#include "ks_dtransfer.hpp"

namespace ksd {

// per-splat constants of the adjoint, built on the host with the splat constants (both precisions)
template <class S> struct SplatAdjT {
    S qn[4];          // normalised quaternion (w, x, y, z) of rows 7-10
    S iqn;            // 1 / |q|
    S dk0k2, dk1k2;   // d(K0/K2)/d ln Theta_e, d(K1/K2)/d ln Theta_e at 1/Theta_e
    S sthB, cthB, sphB, cphB;   // field direction angles
    // derivatives of the per-splat tables with respect to the population rows (x1, x2) = (p, ln gamma_min) for
    // power law, (kappa, ln w) for kappa: power law {dgPL0/dp, dgPL1/dp}; kappa {d2F1/dkappa, d2F1/d ln w,
    // dGamma(k-4/3)/dk, dGamma(k-2)/dk, dGamma(k/4-1/3)/dk, dGamma(k/4+4/3)/dk, dGamma(2+k/2)/dk, d(K0/K2)/d ln w}
    S dpop[8];
    // kappa: the partials of the fits' per-splat factors (kappa_consts) in kappa, ln w and ln B
    S dkk[3][KK_N];
};
using SplatAdj = SplatAdjT<double>;
using SplatAdjf = SplatAdjT<f32>;

// A kappa splat's per-splat factors as five-partial duals over (kappa, ln w, ln B, g, cos B): the values and their
// partials in the three population rows, read as the fits use them
template <class S> struct KappaDualGet {
    const KappaK<S>& k;
    const SplatAdjT<S>& sa;
    KS_DEV DN<5, S> operator()(int f) const {
        DN<5, S> r(k.v[f]);
        r.d[0] = sa.dkk[0][f]; r.d[1] = sa.dkk[1][f]; r.d[2] = sa.dkk[2][f];
        return r;
    }
};

// One sample: given its coefficients, the tail behind it (R = T_{k+1}) and the adjoint w in front of
// it, returns O^T w, j_bar, alpha_bar, rho_bar.
template <class S>
KS_DEV void sample_adjoint(const S j4[4], const S a4[4], const S r3[3], S D, const S w[4], const S R[4],
                           S Otw[4], S jbar[4], S abar[4], S rbar[3]) {
    // q = w . (O R + Ej j) = sum_n o_n w.(K'^n R) + Dp e_n w.(K'^n j). The coefficients o_n, e_n depend on the
    // seven (alpha, rho) only through four invariants (alpha_I, a2, r2, ar): four-partial duals through
    // step_coefs_inv; the products w.(K'^n v) are polynomials in the six entries of K': six-partial duals
    // through three cheap matrix-vector products. The chain rule joins them. Everything is seeded on the
    // Dp-scaled coefficients (O(1) derivatives, no FP32 overflow for near-transparent steps); the bars are
    // scaled back by the power of two Dp at the end, which is exact.
    using D4 = DN<4, S>;
    using D6 = DN<6, S>;
    const S Dp = pow2_below(D);
    const S f = D / Dp;
    S K6v[6];
    const S aIs = a4[0] * Dp;
    for (int i = 0; i < 3; ++i) { K6v[i] = a4[1 + i] * Dp; K6v[3 + i] = r3[i] * Dp; }
    const S a2 = K6v[0] * K6v[0] + K6v[1] * K6v[1] + K6v[2] * K6v[2];
    const S r2 = K6v[3] * K6v[3] + K6v[4] * K6v[4] + K6v[5] * K6v[5];
    const S ar = K6v[0] * K6v[3] + K6v[1] * K6v[4] + K6v[2] * K6v[5];
    D4 o[4], e4[4];
    step_coefs_inv(D4::seed(aIs, 0), D4::seed(a2, 1), D4::seed(r2, 2), D4::seed(ar, 3), f, o, e4);
    // sR[n] = w.(K'^n R), sj[n] = w.(K'^n j)
    D6 K6[6];
    for (int i = 0; i < 6; ++i) K6[i] = D6::seed(K6v[i], i);
    S sR0(0.0), sj0(0.0);
    for (int i = 0; i < 4; ++i) { sR0 += w[i] * R[i]; sj0 += w[i] * j4[i]; }
    D6 vR[4], vj[4], sR[4], sj[4];
    sR[0] = D6(sR0); sj[0] = D6(sj0);
    kprime_mul(K6, R, vR);
    kprime_mul(K6, j4, vj);
    for (int n = 1; n < 4; ++n) {
        sR[n] = vR[0] * w[0] + vR[1] * w[1] + vR[2] * w[2] + vR[3] * w[3];
        sj[n] = vj[0] * w[0] + vj[1] * w[1] + vj[2] * w[2] + vj[3] * w[3];
        if (n < 3) {
            D6 t4[4];
            kprime_mul(K6, vR, t4); for (int i = 0; i < 4; ++i) vR[i] = t4[i];
            kprime_mul(K6, vj, t4); for (int i = 0; i < 4; ++i) vj[i] = t4[i];
        }
    }
    // dq/d(invariant k) and dq/d(K' entry m) at fixed invariants
    S gi[4], hm[6];
    for (int k = 0; k < 4; ++k) {
        S acc(0.0);
        for (int n = 0; n < 4; ++n) acc += o[n].d[k] * sR[n].v + Dp * (e4[n].d[k] * sj[n].v);
        gi[k] = acc;
    }
    for (int m = 0; m < 6; ++m) {
        S acc(0.0);
        for (int n = 1; n < 4; ++n) acc += o[n].v * sR[n].d[m] + Dp * (e4[n].v * sj[n].d[m]);
        hm[m] = acc;
    }
    abar[0] = gi[0] * Dp;
    for (int i = 0; i < 3; ++i) {
        abar[1 + i] = (hm[i] + 2.0 * K6v[i] * gi[1] + K6v[3 + i] * gi[3]) * Dp;        // aQ, aU, aV
        rbar[i] = (hm[3 + i] + 2.0 * K6v[3 + i] * gi[2] + K6v[i] * gi[3]) * Dp;        // rQ, rU, rV
    }
    // values: O^T w = sum o_n K'^T^n w, Ej^T w = Dp sum e_n K'^T^n w
    S u[4], y[4];
    for (int i = 0; i < 4; ++i) { u[i] = w[i]; Otw[i] = o[0].v * w[i]; jbar[i] = Dp * e4[0].v * w[i]; }
    for (int n = 1; n < 4; ++n) {
        kprime_tmul(K6v, u, y);
        for (int i = 0; i < 4; ++i) { u[i] = y[i]; Otw[i] += o[n].v * u[i]; jbar[i] += Dp * e4[n].v * u[i]; }
    }
}

// whether the splat's weight at the sample passes the cutoff, by the same arithmetic as weight_dlog
template <class S>
KS_DEV bool weight_active(const SplatCT<S>& c, S t, S x, S y, S z) {
    S sp, cp;
    if (outside_support(c, t, x, y, z, sp, cp)) return false;
    const S g = gauss_exponent(c, t, x, y, z, sp, cp);
    return !(g > GEXP_SKIP) && exp(-0.5 * g) > WEIGHT_CUTOFF;
}

// rows 7-10 of d ln G: dlnG/dR[a][b] = e_a gu_b (R the rotation of the pattern offset e, gu = dlnG/du) through the
// normalised quaternion (w, x, y, z)
template <class S>
KS_DEV void quat_dlog(const SplatAdjT<S>& sa, const S e[3], const S gu[3], S dl[13]) {
    S GR[9];
    for (int a_ = 0; a_ < 3; ++a_) for (int b = 0; b < 3; ++b) GR[3 * a_ + b] = e[a_] * gu[b];
    const S qw = sa.qn[0], qx = sa.qn[1], qy = sa.qn[2], qz = sa.qn[3];
    S gn[4];
    gn[0] = 2.0 * (-qz * GR[1] + qy * GR[2] + qz * GR[3] - qx * GR[5] - qy * GR[6] + qx * GR[7]);
    gn[1] = 2.0 * (qy * GR[1] + qz * GR[2] + qy * GR[3] - 2.0 * qx * GR[4] - qw * GR[5] + qz * GR[6] + qw * GR[7] - 2.0 * qx * GR[8]);
    gn[2] = 2.0 * (-2.0 * qy * GR[0] + qx * GR[1] + qw * GR[2] + qx * GR[3] + qz * GR[5] - qw * GR[6] + qz * GR[7] - 2.0 * qy * GR[8]);
    gn[3] = 2.0 * (-2.0 * qz * GR[0] - qw * GR[1] + qx * GR[2] + qw * GR[3] - 2.0 * qz * GR[4] + qy * GR[5] + qx * GR[6] + qy * GR[7]);
    const S ng = qw * gn[0] + qx * gn[1] + qy * gn[2] + qz * gn[3];
    for (int i = 0; i < 4; ++i) dl[6 + i] = (gn[i] - sa.qn[i] * ng) * sa.iqn;
}

// whether a splat whose centre follows knots passes the weight cutoff at the sample, by the arithmetic of
// weight_dlog_knot
template <class S>
KS_DEV bool weight_active_knot(const SplatCT<S>& c, const S* knots, const S* ktimes, S t, S x, S y, S z) {
    S cx, cy, cz;
    knot_centre(c, knots, ktimes, t, cx, cy, cz);
    const S e0 = x - cx, e1 = y - cy, e2 = z - cz;
    S q2(0.0);
    for (int b = 0; b < 3; ++b) { const S q = (c.R[b] * e0 + c.R[3 + b] * e1 + c.R[6 + b] * e2) * c.is[b]; q2 += q * q; }
    const S tt = (t - c.t0) * c.iw;
    const S g = q2 + tt * tt;
    return !(g > GEXP_SKIP) && exp(-0.5 * g) > WEIGHT_CUTOFF;
}

// d ln G of a splat whose centre follows knots (trajectories.jl:93-128) with respect to rows 1-12: the centre rows
// shift the whole trajectory (dlnG/dc at the interpolated centre), widths, orientation, t0 and temporal width as for
// pattern splats; the pattern rate does not enter (dl[12] = 0). Returns the Gaussian (0 when culled).
template <class S>
KS_DEV S weight_dlog_knot(const SplatCT<S>& c, const SplatAdjT<S>& sa, const S* knots, const S* ktimes, S t, S x, S y, S z, S dl[13]) {
    S cx, cy, cz;
    knot_centre(c, knots, ktimes, t, cx, cy, cz);
    const S e[3] = {x - cx, y - cy, z - cz};
    S u[3], q[3];
    for (int b = 0; b < 3; ++b) { u[b] = c.R[b] * e[0] + c.R[3 + b] * e[1] + c.R[6 + b] * e[2]; q[b] = u[b] * c.is[b]; }
    const S tt = (t - c.t0) * c.iw;
    const S G = exp(-0.5 * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + tt * tt));
    if (!(G > WEIGHT_CUTOFF)) return S(0.0);
    S gu[3], ge[3];
    for (int b = 0; b < 3; ++b) { gu[b] = -q[b] * c.is[b]; dl[3 + b] = q[b] * q[b]; }
    for (int a_ = 0; a_ < 3; ++a_) ge[a_] = c.R[3 * a_] * gu[0] + c.R[3 * a_ + 1] * gu[1] + c.R[3 * a_ + 2] * gu[2];
    quat_dlog(sa, e, gu, dl);
    for (int a_ = 0; a_ < 3; ++a_) dl[a_] = -ge[a_];
    dl[10] = tt * c.iw;
    dl[11] = tt * tt;
    dl[12] = S(0.0);
    return G;
}

// d ln G of the splat weight (Splats.jl:58-93) with respect to rows 1-12 and the pattern rate; the
// Gaussian itself is returned (0 when culled). dl[0..11] = rows 1..12, dl[12] = omega.
template <class S>
KS_DEV S weight_dlog(const SplatCT<S>& c, const SplatAdjT<S>& sa, S t, S x, S y, S z, S dl[13]) {
    S sp, cp;
    if (outside_support(c, t, x, y, z, sp, cp)) return S(0.0);
    const S cx = c.x * cp - c.y * sp, cy = c.x * sp + c.y * cp;
    const S d0 = x - cx, d1 = y - cy, d2 = z - c.z;
    const S e[3] = {d0 * cp + d1 * sp, -d0 * sp + d1 * cp, d2};
    S u[3], q[3];
    for (int b = 0; b < 3; ++b) { u[b] = c.R[b] * e[0] + c.R[3 + b] * e[1] + c.R[6 + b] * e[2]; q[b] = u[b] * c.is[b]; }
    const S dt = t - c.t0;
    const S tt = dt * c.iw;
    const S gexp = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + tt * tt;
    const S G = exp(-0.5 * gexp);
    if (!(G > WEIGHT_CUTOFF)) return S(0.0);
    // ln G = -(|q|^2 + tt^2) / 2, q_b = u_b exp(-s_b), u = R^T e, e = Rz(-ph) (X - c(ph))
    S gu[3], ge[3];
    for (int b = 0; b < 3; ++b) { gu[b] = -q[b] * c.is[b]; dl[3 + b] = q[b] * q[b]; }
    for (int a_ = 0; a_ < 3; ++a_) ge[a_] = c.R[3 * a_] * gu[0] + c.R[3 * a_ + 1] * gu[1] + c.R[3 * a_ + 2] * gu[2];
    quat_dlog(sa, e, gu, dl);
    // centre: dlnG/dc = -Rz(ph) ge, c = Rz(ph) (c0, c1) + c2 e_z
    const S gd0 = ge[0] * cp - ge[1] * sp, gd1 = ge[0] * sp + ge[1] * cp;
    dl[0] = -(gd0 * cp + gd1 * sp);
    dl[1] = -(gd1 * cp - gd0 * sp);
    dl[2] = -ge[2];
    // pattern phase ph = omega (t - t0): de/dph = (e1 + c1, -(e0 + c0), 0)
    const S dph = ge[0] * (e[1] + c.y) - ge[1] * (e[0] + c.x);
    dl[10] = -c.omega * dph + tt * c.iw;        // t0 (row 11)
    dl[11] = tt * tt;                           // log temporal width (row 12)
    dl[12] = dph * dt;                          // pattern rate
    return G;
}

// The frame of local_frame (g, cos B, chi) and its partials with respect to (theta_B, phi_B, u1, u2, u3):
// the boost of the photon momentum (g, the wave vector k) forward in three-partial duals over u, cos B = k . bh
// from those, and chi by one reverse sweep of its own chain (the polarisation direction f = k x bh, the
// inverse boost, the Walker-Penrose constant, the screen angle) joined to the partials of k and of the boost's
// q = 1 / (gamma + 1). chi depends on the unit direction h = f / |f| only, so the normalisation adds
// nothing to its partials (hbar . h = 0, fbar = hbar / |f|).
template <class S> struct FrameJ { S g, cosB, chi, dg[5], dcosB[5], dchi[5]; };

template <class S>
KS_DEV FrameJ<S> frame_jacobian(const GeoT<S>& G, S a, S alpha, S beta, S thetao_sin, const S uv[3], S sthB, S cthB, S sphB, S cphB) {
    using D3 = DN<3, S>;
    const S bh0 = sthB * cphB, bh1 = sthB * sphB, bh2 = cthB;
    const S bt0 = cthB * cphB, bt1 = cthB * sphB, bt2 = -sthB;     // d bh / d theta_B
    const S bp0 = -bh1, bp1 = bh0;                                  // d bh / d phi_B (third entry 0)
    const D3 u0 = D3::seed(uv[0], 0), u1 = D3::seed(uv[1], 1), u2 = D3::seed(uv[2], 2);
    const D3 gam = sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2));
    const D3 q = 1.0 / (gam + 1.0);
    const D3 ud = u0 * G.pz[1] + u1 * G.pz[2] + u2 * G.pz[3];
    const D3 c1 = ud * q - G.pz[0];
    const D3 g = 1.0 / (gam * G.pz[0] - ud);
    const D3 k0 = (G.pz[1] + u0 * c1) * g, k1 = (G.pz[2] + u1 * c1) * g, k2 = (G.pz[3] + u2 * c1) * g;
    const D3 cosB = k0 * bh0 + k1 * bh1 + k2 * bh2;
    FrameJ<S> o;
    o.g = g.v; o.cosB = cosB.v;
    o.dg[0] = o.dg[1] = S(0.0);
    o.dcosB[0] = k0.v * bt0 + k1.v * bt1 + k2.v * bt2;
    o.dcosB[1] = k0.v * bp0 + k1.v * bp1;
    for (int i = 0; i < 3; ++i) { o.dg[2 + i] = g.d[i]; o.dcosB[2 + i] = cosB.d[i]; }
    for (int i = 0; i < 5; ++i) o.dchi[i] = S(0.0);
    o.chi = S(0.0);
    const S f0 = k1.v * bh2 - k2.v * bh1, f1 = k2.v * bh0 - k0.v * bh2, f2 = k0.v * bh1 - k1.v * bh0;
    const S fn = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
    if (!(fn > 0.0)) return o;
    // forward values (the expressions of local_frame)
    const S ifn = 1.0 / fn;
    const S h0 = f0 * ifn, h1 = f1 * ifn, h2 = f2 * ifn;
    const S uf = u0.v * h0 + u1.v * h1 + u2.v * h2;
    const S ufq = uf * q.v;
    const S fz1 = h0 + u0.v * ufq, fz2 = h1 + u1.v * ufq, fz3 = h2 + u2.v * ufq;
    const S ft = uf * (1.0 / G.lapse);
    const S fr_ = fz1 * (1.0 / G.sgrr);
    const S fth = fz3 * (-1.0 / G.sgthth);
    const S fph = fz2 * (1.0 / G.sgpp) + G.omega * ft;
    const S* p = G.pu;
    const S C1 = a * G.st * G.st, C2 = G.r * G.r + a * a, act = a * G.ct;
    const S A = p[0] * fr_ - p[1] * ft + C1 * (p[1] * fph - p[3] * fr_);
    const S Bw = (C2 * (p[3] * fth - p[2] * fph) - a * (p[0] * fth - p[2] * ft)) * G.st;
    const S kk1 = A * G.r - Bw * act, kk2 = -(A * act + Bw * G.r);
    const S mu = -(alpha + a * thetao_sin);
    const S ea = beta * kk2 - mu * kk1, eb = beta * kk1 + mu * kk2;
    o.chi = atan2(-ea, eb);
    // reverse: d chi = (ea d eb - eb d ea) / (ea^2 + eb^2)
    const S den = ea * ea + eb * eb;
    if (!(den > 0.0)) return o;
    const S w = 1.0 / den;
    const S eab = -eb * w, ebb = ea * w;
    const S kk1b = beta * ebb - mu * eab, kk2b = beta * eab + mu * ebb;
    const S Ab = G.r * kk1b - act * kk2b, Bws = -(act * kk1b + G.r * kk2b) * G.st;
    const S fthb = Bws * (C2 * p[3] - a * p[0]);
    const S fphb = Ab * (C1 * p[1]) - Bws * (C2 * p[2]);
    const S ftb = Bws * (a * p[2]) - Ab * p[1] + fphb * G.omega;
    const S frb = Ab * (p[0] - C1 * p[3]);
    const S z1b = frb * (1.0 / G.sgrr), z2b = fphb * (1.0 / G.sgpp), z3b = fthb * (-1.0 / G.sgthth);
    const S ufqb = z1b * u0.v + z2b * u1.v + z3b * u2.v;
    const S ufb = ftb * (1.0 / G.lapse) + ufqb * q.v;
    const S qb = ufqb * uf;
    const S fb0 = (z1b + ufb * u0.v) * ifn, fb1 = (z2b + ufb * u1.v) * ifn, fb2 = (z3b + ufb * u2.v) * ifn;
    // f = k x bh: kbar = bh x fbar, bhbar = fbar x k
    const S kb0 = bh1 * fb2 - bh2 * fb1, kb1 = bh2 * fb0 - bh0 * fb2, kb2 = bh0 * fb1 - bh1 * fb0;
    const S bb0 = fb1 * k2.v - fb2 * k1.v, bb1 = fb2 * k0.v - fb0 * k2.v, bb2 = fb0 * k1.v - fb1 * k0.v;
    o.dchi[0] = bb0 * bt0 + bb1 * bt1 + bb2 * bt2;
    o.dchi[1] = bb0 * bp0 + bb1 * bp1;
    const S ub[3] = {z1b * ufq + ufb * h0, z2b * ufq + ufb * h1, z3b * ufq + ufb * h2};
    for (int i = 0; i < 3; ++i) o.dchi[2 + i] = ub[i] + qb * q.d[i] + kb0 * k0.d[i] + kb1 * k1.d[i] + kb2 * k2.d[i];
    return o;
}

// One thermal splat's share of one sample (dual_gradient.jl:46-76), given the splat's weight Gw > 0 at
// the sample and d ln G (weight_dlog): adds into grad (the splat's 21-row column).
// bars = (j_bar[4], alpha_bar[4], rho_bar[3]).
// The seven fluid partials (ln Theta_e, ln B, theta_B, phi_B, u1..u3) are joined by the chain rule from the
// frame (g, cos B, chi) with respect to (theta_B, phi_B, u1..u3) (frame_jacobian) and the sample's
// contribution q with respect to (ln Theta_e, ln B, g, cos B) in four-partial duals, with dq/dchi written
// out. The frame enters q only through those three values, so the products are the seven-partial derivatives.
template <class S>
KS_DEV void element_adjoint_thermal(const SplatCT<S>& c, const SplatAdjT<S>& sa, const GeoT<S>& G, S a, S alpha, S beta,
                                    S thetao_sin, S nu, const S bars[11], S Gw, const S dl[13], S grad[21]) {
    const S uv[3] = {c.u[0], c.u[1], c.u[2]};
    const FrameJ<S> fr = frame_jacobian(G, a, alpha, beta, thetao_sin, uv, sa.sthB, sa.cthB, sa.sphB, sa.cphB);
    // the contribution: partials ln Theta_e, ln B, g, cos B; chi enters only through the screen rotation,
    // q = A + Bc cos 2chi + Bs sin 2chi with A, Bc, Bs free of chi, so its derivative is written out
    using D4 = DN<4, S>;
    D4 Th(c.p1); Th.d[0] = c.p1;
    D4 k0k2(c.p2); k0k2.d[0] = sa.dk0k2;
    D4 k1k2(c.p3); k1k2.d[0] = sa.dk1k2;
    D4 Bm(c.Bmag); Bm.d[1] = c.Bmag;
    const D4 g = D4::seed(fr.g, 2), cosB = D4::seed(fr.cosB, 3);
    D4 nuf = nu / g;
    D4 cb = dmin(dmax(cosB, -1.0), 1.0);
    D4 sth, cth;
    pitch_sincos(cb, sth, cth);
    const S ne = exp(c.lnA) * Gw;
    CoefT<D4> cf = thermal(ne, Th, k0k2, k1k2, Bm, nuf, sth, cth);
    if (!(cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ.v != 0.0 || cf.rV.v != 0.0)) return;
    D4 fe = cap_factor(cf.jI, cf.jQ, cf.jV), fa = cap_factor(cf.aI, cf.aQ, cf.aV);
    D4 nh = nuf * (1.0 / NU0);
    D4 i2 = 1.0 / (nh * nh);
    S s2c, c2c;
    sincos(2.0 * fr.chi, &s2c, &c2c);
    D4 jQ = fe * cf.jQ * i2, aQ = fa * cf.aQ * nh, rQ = cf.rQ * nh;
    D4 A = bars[0] * (cf.jI * i2) + bars[3] * (fe * cf.jV * i2) + bars[4] * (cf.aI * nh) + bars[7] * (fa * cf.aV * nh) + bars[10] * (cf.rV * nh);
    D4 Bc = bars[1] * jQ + bars[5] * aQ + bars[8] * rQ;
    D4 Bs = bars[2] * jQ + bars[6] * aQ + bars[9] * rQ;
    D4 q = A + Bc * c2c + Bs * s2c;
    const S qchi = 2.0 * (Bs.v * c2c - Bc.v * s2c);
    const S sv = q.v;
    for (int r = 0; r < 12; ++r) grad[r] += sv * dl[r];
    grad[12] += sv;
    grad[13] += q.d[0];
    grad[14] += q.d[1];
    for (int k = 0; k < 5; ++k) grad[15 + k] += q.d[2] * fr.dg[k] + q.d[3] * fr.dcosB[k] + qchi * fr.dchi[k];
    grad[20] += sv * dl[12];
}

// The X-dependent rotativity factors of kappa_coefs with their derivatives in X (kappa_rhoQ_x, kappa_rhoV_x).
template <class S>
KS_DEV void kappa_rhoQ_x_d(int n, S X, S p084, S& h, S& dh) {
    const double d = n == 1 ? 30.0 : (n == 2 ? 18.0 : (n == 3 ? 12.0 : 8.0));
    const double sd = n == 1 ? 10.0 : (n == 2 ? 6.0 : 4.0);         // sin(X / sd); the fourth fit takes sin(0.375 X)
    const double is_ = n == 4 ? 0.375 : 1.0 / sd;
    const double cc = n == 1 ? 1.5 : (n == 2 ? 1.75 : (n == 3 ? 2.0 : 2.25));
    const double b = n == 1 ? 0.471 : (n == 2 ? 0.5 : (n == 3 ? 0.525 : 0.541));
    const S xb = n == 2 ? sqrt(X) : pow(X, b);
    const S e1 = exp(-p084 / d), E = exp(-cc * xb);
    S sn, cs;
    sincos(n == 4 ? 0.375 * X : X / sd, &sn, &cs);
    h = 1.0 - e1 - sn * E;
    dh = e1 * (0.84 / d) * p084 / X - E * (cs * is_ - sn * (cc * b) * xb / X);
}
template <class S>
KS_DEV void kappa_rhoV_x_d(int n, S xm, S& h, S& dh) {
    const double b = n == 1 ? 0.447 : (n == 2 ? 0.391 : (n == 3 ? 0.348 : 0.313));
    h = 1.0 - 0.17 * log1p(b * xm);
    dh = -0.17 * b / (1.0 + b * xm);
}

// The kappa coefficients of one sample (kappa_coefs) with the reverse sweep of a scalar q of them: the forward
// pass, then qbar(c, cbar), the caller's dq/dc of the coefficient values, then the partials of q in the splat's
// factors (bK, added), the fluid-frame frequency (bnu) and the pitch angle's sine and cosine (bsth, bcth; zero when
// the sine was floored). Where a base of a power is zero its partials are zero (pow_nonneg's choice).
template <class S, class KF, class QBar>
KS_DEV CoefT<S> kappa_coefs_rev(const KF& K, S ne, S nu, S sth, S cth, QBar&& qbar, S bK[KK_N], S& bnu, S& bsth, S& bcth) {
    bool floored = false;
    if (!(sth > 0.0)) { sth = S(sin_floor(sth)); floored = true; }
    const double small = 1e-40;
    // ---- forward (the operations of kappa_coefs)
    const S NUW1 = K(KK_NUW1), JPRE = K(KK_JPRE), APRE = K(KK_APRE);
    const S X = nu / (NUW1 * sth);
    const S pref = ne * JPRE * sth;
    const S X13 = cbrt(X);
    const S lX = log(X);
    const S EI = K(KK_EI), RI = K(KK_RI), MI = K(KK_MI), NLI = K(KK_NLOWI), XI = K(KK_XI);
    const S pXI = pow(X, EI), tI = 1.0 + pXI * RI, PI_ = pow(tI, MI);
    CoefT<S> c;
    c.jI = pref * NLI * X13 * PI_;
    const S EQ = K(KK_EQ), RQ = K(KK_RQ), MQ = K(KK_MQ), NLQ = K(KK_NLOWQ);
    const S pXQ = pow(X, EQ), tQ = 1.0 + pXQ * RQ, PQ = pow(tQ, MQ);
    c.jQ = -(pref * NLQ * X13 * PQ);
    const S c2 = cth * cth;
    const bool near1 = c2 < 0.5;
    const S ls2 = near1 ? log1p(-c2) : 2.0 * log(sth);
    const S u24 = expm1(-1.2 * ls2), u25 = expm1(-1.25 * ls2), u228 = expm1(-1.14 * ls2), u205 = expm1(-1.025 * ls2);
    const S Xm035 = pow(X, -0.35), sX = sqrt(X);
    const S CV1 = K(KK_CV1), CV2 = K(KK_CV2);
    const S v48 = pow_nonneg(u24, 0.48), v44 = pow_nonneg(u25, 0.44);
    const S NlowV = CV1 * v48 * Xm035;
    const S NhighV = CV2 * v44 / sX - small;
    const double sgn = cth > 0.0 ? 1.0 : (cth < 0.0 ? -1.0 : 0.0);
    const bool jv = !(fabs(NhighV) < small * small);
    S rVr(0.0), pV(0.0), tV(1.0), PV(1.0);
    if (jv) {
        rVr = NlowV / NhighV; pV = pow(rVr, XI); tV = 1.0 + pXI * pV; PV = pow(tV, MI);
        c.jV = -(pref * NlowV * X13 * PV) * sgn;
    } else {
        c.jV = S(0.0);
    }
    const S apref = ne * APRE / sth;
    const S Xm53 = pow(X, -5.0 / 3.0);
    const S EA = K(KK_EA), RA = K(KK_RA), MA = K(KK_MA), LA = K(KK_LOWA);
    const S pXA = pow(X, EA), tA = 1.0 + pXA * RA, PA = pow(tA, MA);
    c.aI = apref * LA * Xm53 * PA;
    const S EQA = K(KK_EQA), RQA = K(KK_RQA), MQA = K(KK_MQA), NLQA = K(KK_NLOWQA);
    const S pXQA = pow(X, EQA), tQA = 1.0 + pXQA * RQA, PQA = pow(tQA, MQA);
    c.aQ = -(apref * NLQA * Xm53 * PQA);
    const S CVA1 = K(KK_CVA1), CVA2 = K(KK_CVA2), XVA = K(KK_XVA), EVA = K(KK_EVA), MVA = K(KK_MVA);
    const S v446 = pow_nonneg(u228, 0.446), s205 = sqrt(u205);
    const S NlowVA = CVA1 * v446 * Xm035;
    const S NhighVA = CVA2 * s205 / sX + small;
    const S rVA = NlowVA / NhighVA, pVA = pow(rVA, XVA), pXVA = pow(X, EVA), tVA = 1.0 + pXVA * pVA, PVA = pow(tVA, MVA);
    c.aV = -(apref * NlowVA * Xm53 * PVA) * sgn;
    const int i = (int)val(K(KK_IFIT));
    const S p084 = pow(X, 0.84), xm = 1.0 / sX;
    S hq0, dhq0, hq1, dhq1, hv0, dhv0, hv1, dhv1;
    kappa_rhoQ_x_d<S>(i, X, p084, hq0, dhq0);
    kappa_rhoQ_x_d<S>(i + 1, X, p084, hq1, dhq1);
    kappa_rhoV_x_d<S>(i, xm, hv0, dhv0);
    kappa_rhoV_x_d<S>(i + 1, xm, hv1, dhv1);
    const S FQ0 = K(KK_FQ0), FQ1 = K(KK_FQ1), FV0 = K(KK_FV0), FV1 = K(KK_FV1), RHOQ = K(KK_RHOQ), RHOV = K(KK_RHOV);
    const S fQ = FQ0 * hq0 + FQ1 * hq1, fV = FV0 * hv0 + FV1 * hv1;
    const S inu = 1.0 / nu;
    c.rQ = -(ne * RHOQ * sth * sth) * inu * inu * inu * fQ;
    c.rV = ne * RHOV * cth * inu * inu * fV;
    // ---- the caller's adjoints of the coefficients
    S cb[8];
    qbar(c, cb);
    const S bjI = cb[0], bjQ = cb[1], bjV = cb[2], baI = cb[3], baQ = cb[4], baV = cb[5], brQ = cb[6], brV = cb[7];
    // ---- reverse
    S bX(0.0), bs(0.0), bc(0.0), binu(0.0), bpref(0.0), bapref(0.0), bX13(0.0), bXm53(0.0), bXm035(0.0), bsX(0.0), bls2(0.0), bpXI(0.0);
    {   // rV = ne RHOV cth inu^2 fV, rQ = -ne RHOQ sth^2 inu^3 fQ
        const S t = brV * ne, i2 = inu * inu;
        bK[KK_RHOV] += t * cth * i2 * fV;
        bc += t * RHOV * i2 * fV;
        binu += t * RHOV * cth * 2.0 * inu * fV;
        const S bfV = t * RHOV * cth * i2;
        const S u = -(brQ * ne), i3 = i2 * inu;
        bK[KK_RHOQ] += u * sth * sth * i3 * fQ;
        bs += u * RHOQ * 2.0 * sth * i3 * fQ;
        binu += u * RHOQ * sth * sth * 3.0 * i2 * fQ;
        const S bfQ = u * RHOQ * sth * sth * i3;
        bK[KK_FV0] += bfV * hv0; bK[KK_FV1] += bfV * hv1;
        const S bxm = bfV * (FV0 * dhv0 + FV1 * dhv1);
        bK[KK_FQ0] += bfQ * hq0; bK[KK_FQ1] += bfQ * hq1;
        bX += bfQ * (FQ0 * dhq0 + FQ1 * dhq1);
        bsX += bxm * (-(xm * xm));
    }
    {   // aV = -sgn apref NlowVA Xm53 PVA
        const S g0 = -(baV * sgn);
        bapref += g0 * NlowVA * Xm53 * PVA;
        S bNlowVA = g0 * apref * Xm53 * PVA;
        bXm53 += g0 * apref * NlowVA * PVA;
        const S bPVA = g0 * apref * NlowVA * Xm53;
        const S btVA = bPVA * MVA * PVA / tVA;
        bK[KK_MVA] += bPVA * PVA * log(tVA);
        const S bpXVA = btVA * pVA, bpVA = btVA * pXVA;
        bX += bpXVA * EVA * pXVA / X;
        bK[KK_EVA] += bpXVA * pXVA * lX;
        S brVA(0.0);
        if (rVA > 0.0) { brVA = bpVA * XVA * pVA / rVA; bK[KK_XVA] += bpVA * pVA * log(rVA); }
        bNlowVA += brVA / NhighVA;
        const S bNhighVA = -(brVA * rVA / NhighVA);
        bK[KK_CVA2] += bNhighVA * s205 / sX;
        const S bs205 = bNhighVA * CVA2 / sX;
        bsX += -(bNhighVA * CVA2 * s205 / (sX * sX));
        const S bu205 = s205 > 0.0 ? bs205 * 0.5 / s205 : S(0.0);
        bK[KK_CVA1] += bNlowVA * v446 * Xm035;
        const S bv446 = bNlowVA * CVA1 * Xm035;
        bXm035 += bNlowVA * CVA1 * v446;
        const S bu228 = u228 > 0.0 ? bv446 * 0.446 * v446 / u228 : S(0.0);
        bls2 += bu205 * (-1.025) * (u205 + 1.0) + bu228 * (-1.14) * (u228 + 1.0);
    }
    {   // aQ = -apref NLQA Xm53 PQA, aI = apref LA Xm53 PA
        const S g0 = -baQ;
        bapref += g0 * NLQA * Xm53 * PQA;
        bK[KK_NLOWQA] += g0 * apref * Xm53 * PQA;
        bXm53 += g0 * apref * NLQA * PQA;
        const S bPQA = g0 * apref * NLQA * Xm53;
        const S btQA = bPQA * MQA * PQA / tQA;
        bK[KK_MQA] += bPQA * PQA * log(tQA);
        const S bpXQA = btQA * RQA;
        bK[KK_RQA] += btQA * pXQA;
        bX += bpXQA * EQA * pXQA / X;
        bK[KK_EQA] += bpXQA * pXQA * lX;
        bapref += baI * LA * Xm53 * PA;
        bK[KK_LOWA] += baI * apref * Xm53 * PA;
        bXm53 += baI * apref * LA * PA;
        const S bPA = baI * apref * LA * Xm53;
        const S btA = bPA * MA * PA / tA;
        bK[KK_MA] += bPA * PA * log(tA);
        const S bpXA = btA * RA;
        bK[KK_RA] += btA * pXA;
        bX += bpXA * EA * pXA / X;
        bK[KK_EA] += bpXA * pXA * lX;
    }
    bX += bXm53 * (-5.0 / 3.0) * Xm53 / X;
    bK[KK_APRE] += bapref * ne / sth;
    bs += -(bapref * apref / sth);
    if (jv) {   // jV = -sgn pref NlowV X13 PV
        const S g0 = -(bjV * sgn);
        bpref += g0 * NlowV * X13 * PV;
        S bNlowV = g0 * pref * X13 * PV;
        bX13 += g0 * pref * NlowV * PV;
        const S bPV = g0 * pref * NlowV * X13;
        const S btV = bPV * MI * PV / tV;
        bK[KK_MI] += bPV * PV * log(tV);
        bpXI += btV * pV;
        const S bpV = btV * pXI;
        S brV(0.0);
        if (rVr > 0.0) { brV = bpV * XI * pV / rVr; bK[KK_XI] += bpV * pV * log(rVr); }
        bNlowV += brV / NhighV;
        const S bNhighV = -(brV * rVr / NhighV);
        bK[KK_CV2] += bNhighV * v44 / sX;
        const S bv44 = bNhighV * CV2 / sX;
        bsX += -(bNhighV * CV2 * v44 / (sX * sX));
        const S bu25 = u25 > 0.0 ? bv44 * 0.44 * v44 / u25 : S(0.0);
        bK[KK_CV1] += bNlowV * v48 * Xm035;
        const S bv48 = bNlowV * CV1 * Xm035;
        bXm035 += bNlowV * CV1 * v48;
        const S bu24 = u24 > 0.0 ? bv48 * 0.48 * v48 / u24 : S(0.0);
        bls2 += bu25 * (-1.25) * (u25 + 1.0) + bu24 * (-1.2) * (u24 + 1.0);
    }
    bX += bsX * 0.5 / sX + bXm035 * (-0.35) * Xm035 / X;
    if (near1) bc += bls2 * (-2.0) * cth / (1.0 - c2);
    else bs += bls2 * 2.0 / sth;
    {   // jQ = -pref NLQ X13 PQ, jI = pref NLI X13 PI
        const S g0 = -bjQ;
        bpref += g0 * NLQ * X13 * PQ;
        bK[KK_NLOWQ] += g0 * pref * X13 * PQ;
        bX13 += g0 * pref * NLQ * PQ;
        const S bPQ = g0 * pref * NLQ * X13;
        const S btQ = bPQ * MQ * PQ / tQ;
        bK[KK_MQ] += bPQ * PQ * log(tQ);
        const S bpXQ = btQ * RQ;
        bK[KK_RQ] += btQ * pXQ;
        bX += bpXQ * EQ * pXQ / X;
        bK[KK_EQ] += bpXQ * pXQ * lX;
        bpref += bjI * NLI * X13 * PI_;
        bK[KK_NLOWI] += bjI * pref * X13 * PI_;
        bX13 += bjI * pref * NLI * PI_;
        const S bPI = bjI * pref * NLI * X13;
        const S btI = bPI * MI * PI_ / tI;
        bK[KK_MI] += bPI * PI_ * log(tI);
        bpXI += btI * RI;
        bK[KK_RI] += btI * pXI;
    }
    bX += bpXI * EI * pXI / X;
    bK[KK_EI] += bpXI * pXI * lX;
    bX += bX13 * X13 / (3.0 * X);
    bK[KK_JPRE] += bpref * ne * sth;
    bs += bpref * ne * JPRE;
    bnu = bX * X / nu + binu * (-(inu * inu));
    bK[KK_NUW1] += -(bX * X / NUW1);
    bs += -(bX * X / sth);
    bsth = floored ? S(0.0) : bs;
    bcth = bc;
    return c;
}

// One kappa splat's share of one sample (the element of KappaSplats, populations.jl:34-60) by the reverse sweep of
// its coefficients (kappa_coefs_rev), with the same 22 rows (kappa, ln w the population rows): the screen sums
// and caps give the coefficients' adjoints, the sweep gives q's partials in the splat's factors (joined to kappa,
// ln w and ln B through their partials dkk), in the fluid-frame frequency (to g) and in the pitch angle (to cos B);
// the frame rows join through frame_jacobian. kappa's 2F1 factor is differentiated with kappa and w (its partials
// are among the factors'), not held. The same partials as five-partial forward duals, in fewer operations and
// registers.
template <class S>
KS_DEV void element_adjoint_kappa(const SplatCT<S>& c, const SplatAdjT<S>& sa, const GeoT<S>& G, S a, S alpha, S beta, S thetao_sin,
                                  S nu, const S bars[11], S Gw, const S dl[13], S grad[22]) {
    const S uv[3] = {c.u[0], c.u[1], c.u[2]};
    const FrameJ<S> fr = frame_jacobian(G, a, alpha, beta, thetao_sin, uv, sa.sthB, sa.cthB, sa.sphB, sa.cphB);
    const S nuf = nu / fr.g;
    const bool clipped = !(fr.cosB > -1.0 && fr.cosB < 1.0);
    const S cb = dmin(dmax(fr.cosB, -1.0), 1.0);
    S sth, cth;
    pitch_sincos(cb, sth, cth);
    const S ne = exp(c.lnA) * Gw;
    const S nh = nuf * (1.0 / NU0);
    const S i2 = 1.0 / (nh * nh);
    S s2c, c2c;
    sincos(2.0 * fr.chi, &s2c, &c2c);
    S q(0.0), qchi(0.0), bnh(0.0);
    bool on = true;
    // q = i2 (b0 jI + fe (b3 jV + Bj jQ)) + nh (b4 aI + fa (b7 aV + Ba aQ) + b10 rV + Br rQ), with the caps fe, fa
    auto qbar = [&](const CoefT<S>& cf, S cbar[8]) {
        for (int k = 0; k < 8; ++k) cbar[k] = S(0.0);
        on = cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ != 0.0 || cf.rV != 0.0;
        if (!on) return;
        const S Bj = bars[1] * c2c + bars[2] * s2c, Ba = bars[5] * c2c + bars[6] * s2c, Br = bars[8] * c2c + bars[9] * s2c;
        S fe, fa, dfe[3], dfa[3];
        cap_factor_d(cf.jI, cf.jQ, cf.jV, fe, dfe);
        cap_factor_d(cf.aI, cf.aQ, cf.aV, fa, dfa);
        const S Wj = bars[3] * cf.jV + Bj * cf.jQ, Wa = bars[7] * cf.aV + Ba * cf.aQ;
        const S qj = bars[0] * cf.jI + fe * Wj, qa = bars[4] * cf.aI + fa * Wa + bars[10] * cf.rV + Br * cf.rQ;
        q = i2 * qj + nh * qa;
        cbar[0] = i2 * (bars[0] + Wj * dfe[0]);
        cbar[1] = i2 * (fe * Bj + Wj * dfe[1]);
        cbar[2] = i2 * (fe * bars[3] + Wj * dfe[2]);
        cbar[3] = nh * (bars[4] + Wa * dfa[0]);
        cbar[4] = nh * (fa * Ba + Wa * dfa[1]);
        cbar[5] = nh * (fa * bars[7] + Wa * dfa[2]);
        cbar[6] = nh * Br;
        cbar[7] = nh * bars[10];
        bnh = -2.0 * i2 / nh * qj + qa;
        // d q / d chi: the cos 2chi, sin 2chi parts
        const S Bc = bars[1] * fe * cf.jQ * i2 + bars[5] * fa * cf.aQ * nh + bars[8] * cf.rQ * nh;
        const S Bs = bars[2] * fe * cf.jQ * i2 + bars[6] * fa * cf.aQ * nh + bars[9] * cf.rQ * nh;
        qchi = 2.0 * (Bs * c2c - Bc * s2c);
    };
    S bK[KK_N];
    for (int f = 0; f < KK_N; ++f) bK[f] = S(0.0);
    S bnu, bsth, bcth;
    kappa_coefs_rev(KappaGet<S>{c.kk}, ne, nuf, sth, cth, qbar, bK, bnu, bsth, bcth);
    if (!on) return;
    S dk(0.0), dw(0.0), dB(0.0);
    for (int f = 0; f < KK_N; ++f) { dk += bK[f] * sa.dkk[0][f]; dw += bK[f] * sa.dkk[1][f]; dB += bK[f] * sa.dkk[2][f]; }
    // nu_f = nu / g: d/dg = -(nu_f / g) d/dnu_f; cos B: cth = cb, sth = sqrt((1 - cb)(1 + cb))
    const S dnuf = bnu + bnh * (1.0 / NU0);
    const S dg = -(dnuf * nuf / fr.g);
    const S dcb = clipped ? S(0.0) : bcth + (sth > 0.0 ? bsth * (-(cb / sth)) : S(0.0));
    const S sv = q;
    for (int r = 0; r < 12; ++r) grad[r] += sv * dl[r];
    grad[12] += sv;
    grad[13] += dk;
    grad[14] += dw;
    grad[15] += dB;
    for (int k = 0; k < 5; ++k) grad[16 + k] += dg * fr.dg[k] + dcb * fr.dcosB[k] + qchi * fr.dchi[k];
    grad[21] += sv * dl[12];
}

// One power-law splat's share of one sample (the element of PowerLawSplats, populations.jl:13-33, differentiated as
// element_adjoint_thermal): the 22 rows are the 12 geometric rows, ln n_e, the two population rows (p, ln gamma_min),
// ln B, the field angles, u and the pattern rate. The contribution q is taken in five-partial duals over (p,
// ln gamma_min, ln B, g, cos B); the Gamma products enter with their derivatives in p (dpop); the frame rows join by
// the chain rule through frame_jacobian.
template <class S>
KS_DEV void element_adjoint_pop(const SplatCT<S>& c, const SplatAdjT<S>& sa, const GeoT<S>& G, S a, S alpha, S beta, S thetao_sin,
                                S nu, const S bars[11], S Gw, const S dl[13], S grad[22]) {
    const S uv[3] = {c.u[0], c.u[1], c.u[2]};
    const FrameJ<S> fr = frame_jacobian(G, a, alpha, beta, thetao_sin, uv, sa.sthB, sa.cthB, sa.sphB, sa.cphB);
    using D5 = DN<5, S>;
    D5 Bm(c.Bmag); Bm.d[2] = c.Bmag;
    const D5 g = D5::seed(fr.g, 3), cosB = D5::seed(fr.cosB, 4);
    D5 nuf = nu / g;
    D5 cb = dmin(dmax(cosB, -1.0), 1.0);
    D5 sth, cth;
    pitch_sincos(cb, sth, cth);
    const D5 ne(exp(c.lnA) * Gw);
    CoefT<D5> cf;
    const D5 p = D5::seed(c.p1, 0);
    D5 gmin(c.p2); gmin.d[1] = c.p2;
    D5 gPL[2];
    for (int k = 0; k < 2; ++k) { gPL[k] = D5(c.g[k]); gPL[k].d[0] = sa.dpop[k]; }
    cf = powerlaw(ne, p, gmin, D5(c.p3), gPL, Bm, nuf, sth, cth);
    if (!(cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ.v != 0.0 || cf.rV.v != 0.0)) return;
    D5 fe = cap_factor(cf.jI, cf.jQ, cf.jV), fa = cap_factor(cf.aI, cf.aQ, cf.aV);
    D5 nh = nuf * (1.0 / NU0);
    D5 i2 = 1.0 / (nh * nh);
    S s2c, c2c;
    sincos(2.0 * fr.chi, &s2c, &c2c);
    D5 jQ = fe * cf.jQ * i2, aQ = fa * cf.aQ * nh, rQ = cf.rQ * nh;
    D5 A = bars[0] * (cf.jI * i2) + bars[3] * (fe * cf.jV * i2) + bars[4] * (cf.aI * nh) + bars[7] * (fa * cf.aV * nh) + bars[10] * (cf.rV * nh);
    D5 Bc = bars[1] * jQ + bars[5] * aQ + bars[8] * rQ;
    D5 Bs = bars[2] * jQ + bars[6] * aQ + bars[9] * rQ;
    D5 q = A + Bc * c2c + Bs * s2c;
    const S qchi = 2.0 * (Bs.v * c2c - Bc.v * s2c);
    const S sv = q.v;
    for (int r = 0; r < 12; ++r) grad[r] += sv * dl[r];
    grad[12] += sv;
    grad[13] += q.d[0];
    grad[14] += q.d[1];
    grad[15] += q.d[2];
    for (int k = 0; k < 5; ++k) grad[16 + k] += q.d[3] * fr.dg[k] + q.d[4] * fr.dcosB[k] + qchi * fr.dchi[k];
    grad[21] += sv * dl[12];
}

}  // namespace ksd
