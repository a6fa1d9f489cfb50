// Device Recurrence(M) marcher: the fused march of one ray with a consumer folded over its samples.
//
// Reference: march_ray (src/Geodesics/quadrature.jl:231-296) and the pieces it uses
// (recurrence.jl:24-308, quadrature.jl:42-296, jacobi.jl:8-71). Per sample: two addition-theorem
// steps of the radial and polar Jacobi states to the midpoint and the end of the interval, the
// radius and polar angle from the states, Simpson on the desingularised Mino-time rates of t~ and
// phi plus their closed-form singular pieces (large r, both horizon poles, the polar-axis spike),
// the split at the radial turning point of scattering rays, and every Man samples (M, fewer near
// the critical curve) a re-anchoring of the states and of t~, phi to the closed forms.
#pragma once
#include "ks_dgeo.hpp"

namespace ksd {

struct JS { double sn, cn, dn; };
KS_DEV JS js_at(double u, double m) { Jac j = jac_pos(u, m); return {j.sn, j.cn, j.dn}; }
KS_DEV JS js_step(const JS& x, const JS& d, double m) {
    double den = 1.0 / (1.0 - m * (x.sn * x.sn) * (d.sn * d.sn));
    return {(x.sn * d.cn * d.dn + d.sn * x.cn * x.dn) * den,
            (x.cn * d.cn - x.sn * x.dn * d.sn * d.dn) * den,
            (x.dn * d.dn - m * x.sn * x.cn * d.sn * d.cn) * den};
}
KS_DEV JS js_step_constants(double delta, double m) {
    JS x = js_at(fabs(delta), m);
    return {copysign(x.sn, delta), x.cn, x.dn};
}

struct QC {   // quadrature constants (quadrature.jl:42-97)
    double sp, sm, Qt, Qtm, Qph, Qphm, C, nq, sqrt_eps, gc, mpp;
    bool scattering;
};
KS_DEV QC make_qc(const PixC& p, const Spin& S) {
    QC q;
    const double a = S.a, lam = p.lam;
    double E = S.rp * S.rp + a * a - a * lam, Em = S.rm * S.rm + a * a - a * lam;
    double dr = S.rp - S.rm;
    q.sp = -fabs(E); q.sm = -fabs(Em);
    q.Qt = (S.rp * S.rp + a * a) * E / dr;
    q.Qtm = -(S.rm * S.rm + a * a) * Em / dr;
    q.Qph = a * (2.0 * S.rp - a * lam) / dr;
    q.Qphm = -a * (2.0 * S.rm - a * lam) / dr;
    q.scattering = p.rcase == 2 && !isfinite(p.tau_h) && p.tau_total == 2.0 * p.fo;
    if (p.vortical) { q.C = 1.0 / (1.0 - p.um); q.nq = (p.up - p.um) / (1.0 - p.um); }
    else { q.C = 1.0; q.nq = p.up; }
    q.sqrt_eps = sqrt(dmax(1.0 - q.nq, 0.0));
    q.gc = 1.0 / sqrt(1.0 - p.mth);
    q.mpp = p.mth / (1.0 - p.mth);
    return q;
}

struct RP { double r, th, cth, s, sn, cn, gt, gph; bool nur, nuth; };

KS_DEV void smooth_integrands(const PixC& p, const Spin& S, const QC& q, double r, double cth, double s, double sn, double cn, double dn,
                              double& gt, double& gph) {
    const double a = S.a, lam = p.lam;
    double D = r * r - 2.0 * r + a * a;
    double s2 = 1.0 - cth * cth;
    double iD = 1.0 / D;
    double ft = (r * r + a * a) * (r * r + a * a - a * lam) * iD + a * (lam - a * s2);
    double fph = a * (2.0 * r - a * lam) * iD;
    double wp = s * S.rp / (r * (r - S.rp)) / q.sp;
    double wm = s * S.rm / (r * (r - S.rm)) / q.sm;
    gt = ft - fabs(s) * (1.0 + 2.0 / r) - q.Qt * wp - q.Qtm * wm;
    double gpol = lam * q.C * (1.0 - q.gc * dn * (1.0 - q.mpp * cn * cn * 0.5)) / (1.0 - q.nq * sn * sn);
    gph = fph - q.Qph * wp - q.Qphm * wm + gpol;
}

KS_DEV RP ray_point(const PixC& p, const Spin& S, const QC& q, const JS& xr, const JS& xt, double tau) {
    RP o;
    o.r = radius_jac(p, xr.sn, xr.cn);
    polar_from_jac(p, xt.sn, xt.cn, xt.dn, o.th, o.cth, o.nuth);
    o.nur = before_turning<double>(tau, p.fo);
    double R = r_potential(S.a, p.lam, p.eta, o.r);
    const double w = o.r * o.r + S.a * S.a - S.a * p.lam;
    // R below 1e-14 w^2 is rounding (a sample on the radial turning point, where R = 0): s = 0 there
    o.s = (o.nur ? -1.0 : 1.0) * (R > 1e-14 * (w * w) ? sqrt(R) : 0.0);
    double idn = 1.0 / xt.dn;
    o.sn = xt.sn * idn / p.scale;
    o.cn = xt.cn * idn;
    smooth_integrands(p, S, q, o.r, o.cth, o.s, o.sn, o.cn, idn, o.gt, o.gph);
    return o;
}
KS_COLD JS radial_state(const PixC& p, double tau) { return js_at(p.c * (p.fo - tau), p.k); }
KS_COLD JS polar_state(const PixC& p, double tau) { return js_at(p.scale * (tau + p.off) / p.f, p.mu); }

KS_COLD RP ray_point_turning(const PixC& p, const Spin& S, const QC& q) {
    JS xt = polar_state(p, p.fo);
    RP o;
    o.r = p.r4;
    polar_from_jac(p, xt.sn, xt.cn, xt.dn, o.th, o.cth, o.nuth);
    o.nur = true;
    o.s = 0.0;
    double idn = 1.0 / xt.dn;
    o.sn = xt.sn * idn / p.scale;
    o.cn = xt.cn * idn;
    smooth_integrands(p, S, q, o.r, o.cth, o.s, o.sn, o.cn, idn, o.gt, o.gph);
    return o;
}

KS_DEV double polar_closed_form(const PixC& p, const QC& q, double tau, double sn, double cn) {
    double X = (tau + p.off) / p.f;
    double nw = floor((X + p.K2 * 0.5) / p.K2);
    double A = (atan(q.sqrt_eps * sn / cn) + nw * PI) / q.sqrt_eps;
    double ph = atan(sn / cn) + nw * PI;
    double eps = q.sqrt_eps * q.sqrt_eps;
    return p.lam * q.C * p.f * q.gc * ((1.0 + q.mpp * eps / (2.0 * q.nq)) * A - q.mpp / (2.0 * q.nq) * ph);
}

// Pa, Pb: polar_closed_form at the interval's ends (each end's value serves both intervals that share it)
KS_DEV void interval_increment(const PixC& p, const Spin& S, const QC& q, const RP& pa, const RP& pm, const RP& pb, double ta, double tb,
                               double Pa, double Pb, double& It, double& Iph) {
    double h = tb - ta;
    double ra = pa.r, rb = pb.r;
    double dFr = (rb - ra) + 2.0 * log(rb / ra);
    double dFp = log((rb - S.rp) * ra / ((ra - S.rp) * rb)) / q.sp;
    double dFm = log((rb - S.rm) * ra / ((ra - S.rm) * rb)) / q.sm;
    It = h * (1.0 / 6.0) * (pa.gt + 4.0 * pm.gt + pb.gt) + fabs(dFr) + q.Qt * dFp + q.Qtm * dFm;
    Iph = h * (1.0 / 6.0) * (pa.gph + 4.0 * pm.gph + pb.gph) + q.Qph * dFp + q.Qphm * dFm + Pb - Pa;
}

// the interval holding the radial turning point of a scattering ray, split there (quadrature.jl:255-270)
KS_COLD void turning_increment(const PixC& p, const Spin& S, const QC& q, const RP& pa, const RP& pb, double ta, double tb, double Pa, double Pb,
                               double& t, double& ph) {
    RP pt = ray_point_turning(p, S, q);
    const double Pt = polar_closed_form(p, q, p.fo, pt.sn, pt.cn);
    double t1 = 0.5 * (ta + p.fo), t2 = 0.5 * (p.fo + tb);
    RP p1 = ray_point(p, S, q, radial_state(p, t1), polar_state(p, t1), t1);
    RP p2 = ray_point(p, S, q, radial_state(p, t2), polar_state(p, t2), t2);
    double It1, Ip1, It2, Ip2;
    interval_increment(p, S, q, pa, p1, pt, ta, p.fo, Pa, Pt, It1, Ip1);
    interval_increment(p, S, q, pt, p2, pb, p.fo, tb, Pt, Pb, It2, Ip2);
    t += It1 + It2;
    ph += Ip1 + Ip2;
}

KS_DEV int anchor_interval(int M, double one_minus_k) {
    double scaled = M * sqrt(dmax(one_minus_k, 0.0) / 2e-5);
    int s = (int)dmin(scaled, (double)M);
    int r = M < s ? M : s;
    return r > 4 ? r : 4;
}

// Samples kbeg..kend of one ray; f(k, Samp, dtau) is called for those k only. The largest anchor
// residuals are taken over the anchors k with kbeg < k <= kend + 1 (so that the segments of a ray
// [1, L], [L+1, 2L], ... together cover every anchor exactly once).
//
// A ray can be cut at any anchor: there the Jacobi states are reset to the closed forms, and t~, phi
// too whenever the anchor sample is a valid Krang sample (quadrature.jl:283-291). The march state at
// such an anchor is a function of the pixel alone, so a march started there reproduces the march of
// the whole ray. kbeg must be 1 or an anchor (kbeg - 1 a multiple of the anchor interval); when the
// anchor kbeg is not a valid sample, t~ and phi carry the quadrature from the last valid anchor
// before it, and the march starts from that anchor (or from sample 1, always a closed-form start)
// and runs without calling f up to kbeg.
template <class F>
KS_DEV void march_range(const PixC& p, const Spin& S, double thetao, int N, int M, int kbeg, int kend, F& f, double& resid_t, double& resid_ph) {
    const QC q = make_qc(p, S);
    const VisC vis = vis_constants(p, S.a, thetao);
    const double dtau = p.tau_total / (double)(N + 1);
    const double h = 0.5 * dtau;
    // the radial turning point splits the interval that holds it (ta < fo < tb). With N + 1 even, fo = tau_total / 2
    // falls on a sample in exact arithmetic, and the strict comparison would be decided by the rounding of fo and
    // k dtau; a turning point within 1e-9 dtau of a sample is taken as on it (no split, the exact-arithmetic case)
    const double tol = 1e-9 * dtau;
    const int Man = anchor_interval(M, 1.0 - p.k);
    const JS Dr = js_step_constants(-p.c * h, p.k);
    const JS Dt = js_step_constants(p.scale * h / p.f, p.mu);
    resid_t = 0.0; resid_ph = 0.0;
    // start: the last anchor A <= kbeg whose sample is valid, or sample 1
    int A = kbeg;
    JS xr, xt;
    RP pa;
    for (;;) {
        const double ta = (double)A * dtau;
        xr = radial_state(p, ta);
        xt = polar_state(p, ta);
        pa = ray_point(p, S, q, xr, xt, ta);
        if (A <= 1 || ((ta <= p.tau_h) && krang_visible(vis, pa.cth))) break;
        A = A - Man > 1 ? A - Man : 1;
    }
    double t, ph;
    t_phi_direct(p, S, (double)A * dtau, t, ph);
    double Pa = polar_closed_form(p, q, (double)A * dtau, pa.sn, pa.cn);
    if (A == kbeg) {
        const double ta = (double)A * dtau;
        Samp s{t, pa.r, pa.th, ph, pa.nur, pa.nuth, (ta <= p.tau_h) && krang_visible(vis, pa.cth)};
        f(A, s, dtau);
    }
    const int klast = kend < N ? kend + 1 : N;
    for (int k = A + 1; k <= klast; ++k) {
        const double ta = (double)(k - 1) * dtau, tb = (double)k * dtau;
        const bool anchor = ((k - 1) % Man) == 0;
        if (k > kend && !anchor) break;            // the step past kend only serves the next anchor's residual
        xr = js_step(xr, Dr, p.k);
        xt = js_step(xt, Dt, p.mu);
        RP pmid = ray_point(p, S, q, xr, xt, ta + h);
        if (anchor) { xr = radial_state(p, tb); xt = polar_state(p, tb); }
        else { xr = js_step(xr, Dr, p.k); xt = js_step(xt, Dt, p.mu); }
        RP pb = ray_point(p, S, q, xr, xt, tb);
        const double Pb = polar_closed_form(p, q, tb, pb.sn, pb.cn);
        if (q.scattering && p.fo - ta > tol && tb - p.fo > tol) {
            turning_increment(p, S, q, pa, pb, ta, tb, Pa, Pb, t, ph);
        } else {
            double It, Ip;
            interval_increment(p, S, q, pa, pmid, pb, ta, tb, Pa, Pb, It, Ip);
            t += It;
            ph += Ip;
        }
        const bool okb = (tb <= p.tau_h) && krang_visible(vis, pb.cth);
        if (anchor && okb) {
            // closed-form anchor; Krang's ok for the sample decides whether it resets (quadrature.jl:283-291)
            double td, pd;
            t_phi_direct(p, S, tb, td, pd);
            if (k > kbeg) {
                resid_t = dmax(resid_t, fabs(t - td));
                resid_ph = dmax(resid_ph, fabs(ph - pd));
            }
            t = td; ph = pd;
        }
        if (k >= kbeg && k <= kend) {
            Samp s{t, pb.r, pb.th, ph, pb.nur, pb.nuth, okb};
            f(k, s, dtau);
        }
        pa = pb;
        Pa = Pb;
    }
}

// March one whole ray (samples 1..N). Returns the largest anchor residuals.
template <class F>
KS_DEV void march_ray(const PixC& p, const Spin& S, double thetao, int N, int M, F& f, double& resid_t, double& resid_ph) {
    march_range(p, S, thetao, N, M, 1, N, f, resid_t, resid_ph);
}

// Segments of a ray for parallel marching: whole multiples of the anchor interval, at least SEG_MIN
// samples (at most 2 SEG_MIN - 1, or the anchor interval itself when that is longer).
constexpr int SEG_MIN = 32;
KS_DEV int segment_length(int M, double one_minus_k) {
    const int Man = anchor_interval(M, one_minus_k);
    return Man * ((SEG_MIN + Man - 1) / Man);
}
__host__ __device__ inline int max_segments(int N) { return (N + SEG_MIN - 1) / SEG_MIN; }

}  // namespace ksd
