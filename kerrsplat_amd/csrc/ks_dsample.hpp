// One sample's splat sum and step in the transfer's scalar (double or f32): the context of an item, the
// knot-interpolated Gaussian exponent, the per-sample coefficient sum over a candidate list, the step applied to a
// segment map, and the path weight. Shared by the transfer kernels and the host-side work model.
// Reference: transport.jl:254-290, polarized.jl:110-150, populations.jl:13-104, trajectories.jl:93-128.
#pragma once
#include "ks_kernels.hpp"

namespace ksd {

// ---- splat sums ----------------------------------------------------------------------------------------------
// What a sample's splat sum needs besides the sample itself, in the transfer's scalar S (double or
// f32); the geometry of a sample (metric, tetrad, photon momentum) is always built in FP64.
template <class S>
struct TransCtxT {
    const SplatCT<S>* __restrict__ sp; const S* __restrict__ knots; const S* __restrict__ ktimes;
    int ns, nknots;
    double a, rp, rm, eta, lam, t_obs;          // FP64: geometry and emission times
    S aS, alpha, beta, thetao_sin, nu, L;
};
using TransCtx = TransCtxT<double>;
template <class S>
KS_DEV TransCtxT<S> make_ctx(double a, double rp, double rm, double thetao_sin, double alpha, double beta, double eta, double lam,
                             const SplatCT<S>* sp, int ns, const S* knots, const S* ktimes, int nknots, double t_obs, double nu, double L) {
    TransCtxT<S> c;
    c.sp = sp; c.knots = knots; c.ktimes = ktimes; c.ns = ns; c.nknots = nknots;
    c.a = a; c.rp = rp; c.rm = rm; c.eta = eta; c.lam = lam; c.t_obs = t_obs;
    c.aS = S(a); c.alpha = S(alpha); c.beta = S(beta); c.thetao_sin = S(thetao_sin); c.nu = S(nu); c.L = S(L);
    return c;
}

// The Gaussian exponent at (t, x, y, z) of a splat whose centre follows knots (trajectories.jl:93-128): the offset
// from the knot-interpolated centre (knot_centre), no pattern rotation.
template <class S>
KS_DEV S knot_gexp(const TransCtxT<S>& c, const SplatCT<S>& sc, S t, S x, S y, S z) {
    S cx, cy, cz;
    knot_centre(sc, c.knots, c.ktimes, t, cx, cy, cz);
    S d0 = x - cx, d1 = y - cy, d2 = z - cz;
    S u0 = sc.R[0] * d0 + sc.R[3] * d1 + sc.R[6] * d2;
    S u1 = sc.R[1] * d0 + sc.R[4] * d1 + sc.R[7] * d2;
    S u2 = sc.R[2] * d0 + sc.R[5] * d1 + sc.R[8] * d2;
    S q0 = u0 * sc.is[0], q1 = u1 * sc.is[1], q2 = u2 * sc.is[2];
    S tt = (t - sc.t0) * sc.iw;
    return q0 * q0 + q1 * q1 + q2 * q2 + tt * tt;
}

// The splat sum of one sample (accumulate_elements, transport.jl:254-276): invariant j, alpha, rho in
// screen axes; false when no splat contributes. One instantiation per population and scalar.
// With a candidate list (ascending splat indices, the per-tile cull) only those splats are visited; the
// others lie outside their support at every sample of the tile and would add exact zeros. With gout the sample's
// geometry (in S) is returned when the sample contributes.
template <int KIND, class S>
KS_DEV bool sample_coefficients(const TransCtxT<S>& c, double st, double r, double th, bool nur, bool nuth, S x, S y, S z,
                                S j4[4], S a4[4], S r3[3], const uint16_t* cand = nullptr, int ncand = 0, GeoT<S>* gout = nullptr) {
    const S t = S(c.t_obs - st);
    for (int i = 0; i < 4; ++i) { j4[i] = S(0.0); a4[i] = S(0.0); }
    r3[0] = r3[1] = r3[2] = S(0.0);
    bool active = false, geo_ready = false;
    GeoT<S> G;
    const int nloop = cand ? ncand : c.ns;
    for (int ci = 0; ci < nloop; ++ci) {
        const int i = cand ? (int)cand[ci] : ci;
        const SplatCT<S>& sc = c.sp[i];
        S gexp;
        if (KIND == MODEL_KNOT || sc.nk > 0) {
            gexp = knot_gexp(c, sc, t, x, y, z);
        } else {
            S sph, cph;
            if (outside_support(sc, t, x, y, z, sph, cph)) continue;
            gexp = gauss_exponent(sc, t, x, y, z, sph, cph);
        }
        if (gexp > GEXP_SKIP) continue;
        S G_ = exp(-0.5 * gexp);
        if (!(G_ > WEIGHT_CUTOFF)) continue;
        if (!geo_ready) {
            Geo Gd;
            make_geo(c.a, r, th, c.eta, c.lam, nur, nuth, Gd);
            G = geo_as<S>(Gd);
            geo_ready = true;
        }
        FrameT<S> fr = local_frame(G, c.aS, c.alpha, c.beta, c.thetao_sin, sc.u, sc.bh);
        S nuf = c.nu / fr.g;
        S cb = dmin(dmax(fr.cosB, -1.0), 1.0);
        S sth, cth;
        // FP64 takes cos = cb and the product form of the sine; the FP32 path keeps acos and sincos: with the
        // product form, or with the sine taken in FP64 from cb, its images moved further from the FP64 images
        // (9.0e-7 against 6.2e-7 of peak I, 16 x 16 pixels, N = 256, 4 thermal splats), which the sine's
        // accuracy alone does not explain
        if constexpr (std::is_same<S, double>::value) pitch_sincos(cb, sth, cth);
        else sincos(acos(cb), &sth, &cth);
        S ne = exp(sc.lnA) * G_;
        CoefT<S> cf;
        if constexpr (KIND == MODEL_THERMAL || KIND == MODEL_KNOT) cf = thermal(ne, sc.p1, sc.p2, sc.p3, sc.Bmag, nuf, sth, cth);
        else if constexpr (KIND == MODEL_POWERLAW) cf = powerlaw(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        else if constexpr (KIND == MODEL_KAPPA) cf = kappa_coefs(KappaGet<S>{sc.kk}, ne, nuf, sth, cth);
        else if (sc.kind == MODEL_THERMAL || sc.kind == MODEL_KNOT) cf = thermal(ne, sc.p1, sc.p2, sc.p3, sc.Bmag, nuf, sth, cth);
        else if (sc.kind == MODEL_POWERLAW) cf = powerlaw(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        else cf = kappa_coefs(KappaGet<S>{sc.kk}, ne, nuf, sth, cth);
        if (!(cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ != 0.0 || cf.rV != 0.0)) continue;
        S fe = cap_factor(cf.jI, cf.jQ, cf.jV), fa = cap_factor(cf.aI, cf.aQ, cf.aV);
        S nh = nuf * (1.0 / NU0);
        S i2 = 1.0 / (nh * nh);
        S s2c, c2c; sincos(2.0 * fr.chi, &s2c, &c2c);
        S jQ = fe * cf.jQ * i2, aQ = fa * cf.aQ * nh, rQ = cf.rQ * nh;
        j4[0] += cf.jI * i2; j4[1] += jQ * c2c; j4[2] += jQ * s2c; j4[3] += fe * cf.jV * i2;
        a4[0] += cf.aI * nh; a4[1] += aQ * c2c; a4[2] += aQ * s2c; a4[3] += fa * cf.aV * nh;
        r3[0] += rQ * c2c; r3[1] += rQ * s2c; r3[2] += cf.rV * nh;
        active = true;
    }
    if (gout && active) *gout = G;
    return active;
}

// thin splats (Splats.jl:113-128): sum_i j_i g^2 Sigma dtau with the ZAMO redshift g
KS_DEV double thin_sample(const SplatC* __restrict__ sp, int ns, double a, double lam, double t_obs, double st, double r, double th,
                          double x, double y, double z, double dtau) {
    const double t = t_obs - st;
    double jt = 0.0;
    for (int i = 0; i < ns; ++i) {
        const SplatC& c = sp[i];
        double sph, cph; sincos(c.omega * (t - c.t0), &sph, &cph);
        jt += exp(c.lnA - 0.5 * gauss_exponent(c, t, x, y, z, sph, cph));
    }
    double s_, c_; sincos(th, &s_, &c_);
    double S_ = r * r + a * a * c_ * c_, Dl = r * r - 2.0 * r + a * a;
    double Xi = (r * r + a * a) * (r * r + a * a) - a * a * Dl * s_ * s_;
    double g = sqrt(Dl * S_ / Xi) / (1.0 - 2.0 * a * r / Xi * lam);
    return g * g * jt * S_ * dtau;
}


// The FP64 step of an FP32 sample, out of line (the rare samples with a Faraday depth |rho| Delta above
// FARADAY_FP64, where the step's polynomial in K' cancels strongly): O and E in FP64, rounded to FP32. The
// per-sample sums are passed by address here: the transfer measured 13% slower with them passed by value (Step32).
KS_COLD void step_f64_for_f32(const f32 j4[4], const f32 a4[4], const f32 r3[3], f32 D, f32 O[16], f32 E[4]) {
    double jd[4], ad[4], rd[3], Od[16], Ed[4];
    for (int i = 0; i < 4; ++i) { jd[i] = val(j4[i]); ad[i] = val(a4[i]); }
    for (int i = 0; i < 3; ++i) rd[i] = val(r3[i]);
    transfer_step(jd, ad, rd, val(D), Od, Ed);
    for (int i = 0; i < 16; ++i) O[i] = f32(Od[i]);
    for (int i = 0; i < 4; ++i) E[i] = f32(Ed[i]);
}
// One sample's step applied to a segment map, in the transfer's scalar.
template <class S>
KS_DEV void map_step(SegMapT<S>& sm, const S j4[4], const S a4[4], const S r3[3], S D) {
    if constexpr (!std::is_same<S, double>::value) {
        const S fd = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]) * D;
        if (fd > FARADAY_FP64) {
            S O[16], E[4];
            step_f64_for_f32(j4, a4, r3, D, O, E);
            sm.advance(O, E);
            return;
        }
    }
    sm.step(j4, a4, r3, D);
}

// Delta = (L / nu-hat) Sigma dtau (transport.jl:282-284)
KS_DEV double path_weight(double L, double nu, double a, double r, double th, double dtau) {
    double ct = cos(th);
    return L / (nu * (1.0 / NU0)) * (r * r + a * a * ct * ct) * dtau;
}

}  // namespace ksd
