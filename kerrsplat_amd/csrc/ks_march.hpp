// Device Recurrence(M) marcher: the fused march of one ray with a consumer folded over its samples.
//
// Reference: march_ray (src/Geodesics/quadrature.jl:231-296) and the pieces it uses
// (recurrence.jl:24-517, quadrature.jl:42-215, jacobi.jl:146-209). Per sample: two addition-theorem
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
    o.nur = tau < p.fo;
    double R = r_potential(S.a, p.lam, p.eta, o.r);
    o.s = (o.nur ? -1.0 : 1.0) * sqrt(dmax(R, 0.0));
    double idn = 1.0 / xt.dn;
    o.sn = xt.sn * idn / p.scale;
    o.cn = xt.cn * idn;
    smooth_integrands(p, S, q, o.r, o.cth, o.s, o.sn, o.cn, idn, o.gt, o.gph);
    return o;
}
KS_DEV JS radial_state(const PixC& p, double tau) { return js_at(p.c * (p.fo - tau), p.k); }
KS_DEV JS polar_state(const PixC& p, double tau) { return js_at(p.scale * (tau + p.off) / p.f, p.mu); }

KS_DEV RP ray_point_turning(const PixC& p, const Spin& S, const QC& q) {
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

KS_DEV void interval_increment(const PixC& p, const Spin& S, const QC& q, const RP& pa, const RP& pm, const RP& pb, double ta, double tb,
                               double& It, double& Iph) {
    double h = tb - ta;
    double ra = pa.r, rb = pb.r;
    double dFr = (rb - ra) + 2.0 * log(rb / ra);
    double dFp = log((rb - S.rp) * ra / ((ra - S.rp) * rb)) / q.sp;
    double dFm = log((rb - S.rm) * ra / ((ra - S.rm) * rb)) / q.sm;
    It = h * (1.0 / 6.0) * (pa.gt + 4.0 * pm.gt + pb.gt) + fabs(dFr) + q.Qt * dFp + q.Qtm * dFm;
    Iph = h * (1.0 / 6.0) * (pa.gph + 4.0 * pm.gph + pb.gph) + q.Qph * dFp + q.Qphm * dFm +
          polar_closed_form(p, q, tb, pb.sn, pb.cn) - polar_closed_form(p, q, ta, pa.sn, pa.cn);
}

KS_DEV int anchor_interval(int M, double one_minus_k) {
    double scaled = M * sqrt(dmax(one_minus_k, 0.0) / 2e-5);
    int s = (int)dmin(scaled, (double)M);
    int r = M < s ? M : s;
    return r > 4 ? r : 4;
}

// March one ray; f(k, Samp, dtau) is called for k = 1..N. Returns the largest anchor residuals.
template <class F>
KS_DEV void march_ray(const PixC& p, const Spin& S, double thetao, int N, int M, F& f, double& resid_t, double& resid_ph) {
    const QC q = make_qc(p, S);
    const VisC vis = vis_constants(p, S.a, thetao);
    const double dtau = p.tau_total / (double)(N + 1);
    const double h = 0.5 * dtau;
    const int Man = anchor_interval(M, 1.0 - p.k);
    const JS Dr = js_step_constants(-p.c * h, p.k);
    const JS Dt = js_step_constants(p.scale * h / p.f, p.mu);
    JS xr = radial_state(p, dtau);
    JS xt = polar_state(p, dtau);
    RP pa = ray_point(p, S, q, xr, xt, dtau);
    double t, ph;
    t_phi_direct(p, S, dtau, t, ph);
    resid_t = 0.0; resid_ph = 0.0;
    {
        Samp s{t, pa.r, pa.th, ph, pa.nur, pa.nuth, (dtau <= p.tau_h) && krang_visible(vis, pa.cth)};
        f(1, s, dtau);
    }
    for (int k = 2; k <= N; ++k) {
        const double ta = (double)(k - 1) * dtau, tb = (double)k * dtau;
        xr = js_step(xr, Dr, p.k);
        xt = js_step(xt, Dt, p.mu);
        RP pmid = ray_point(p, S, q, xr, xt, ta + h);
        const bool anchor = ((k - 1) % Man) == 0;
        if (anchor) { xr = radial_state(p, tb); xt = polar_state(p, tb); }
        else { xr = js_step(xr, Dr, p.k); xt = js_step(xt, Dt, p.mu); }
        RP pb = ray_point(p, S, q, xr, xt, tb);
        if (q.scattering && ta < p.fo && p.fo < tb) {
            RP pt = ray_point_turning(p, S, q);
            double t1 = 0.5 * (ta + p.fo), t2 = 0.5 * (p.fo + tb);
            RP p1 = ray_point(p, S, q, radial_state(p, t1), polar_state(p, t1), t1);
            RP p2 = ray_point(p, S, q, radial_state(p, t2), polar_state(p, t2), t2);
            double It1, Ip1, It2, Ip2;
            interval_increment(p, S, q, pa, p1, pt, ta, p.fo, It1, Ip1);
            interval_increment(p, S, q, pt, p2, pb, p.fo, tb, It2, Ip2);
            t += It1 + It2;
            ph += Ip1 + Ip2;
        } else {
            double It, Ip;
            interval_increment(p, S, q, pa, pmid, pb, ta, tb, It, Ip);
            t += It;
            ph += Ip;
        }
        if (anchor) {
            // closed-form anchor; Krang's ok for the sample decides whether it resets (quadrature.jl:283-291)
            double td, pd;
            t_phi_direct(p, S, tb, td, pd);
            bool okd = (tb <= p.tau_h) && krang_visible(vis, pb.cth);
            if (okd) {
                resid_t = dmax(resid_t, fabs(t - td));
                resid_ph = dmax(resid_ph, fabs(ph - pd));
                t = td; ph = pd;
            }
        }
        Samp s{t, pb.r, pb.th, ph, pb.nur, pb.nuth, (tb <= p.tau_h) && krang_visible(vis, pb.cth)};
        f(k, s, dtau);
        pa = pb;
    }
}

}  // namespace ksd
