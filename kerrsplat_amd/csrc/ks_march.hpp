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
#include "ks_ddual.hpp"

namespace ksd {

// The march is written once as templates on its bulk scalar B: double (the plain FP64 march) or f32 (the bulk of
// the per-sample work in FP32: the Jacobi addition-theorem steps, the radius and polar angle, the integrand rates
// and the Simpson increments). The closed-form anchors (Jacobi functions, t~ and phi), the quadrature constants, the
// visibility constants, the Mino times and the running t~ / phi sums stay FP64 for both.
template <class B> struct JST { B sn, cn, dn; };
using JS = JST<double>;
KS_DEV JS js_at(double u, double m) { Jac j = jac_pos(u, m); return {j.sn, j.cn, j.dn}; }
template <class B>
KS_DEV JST<B> js_step(const JST<B>& x, const JST<B>& d, B m) {
    B den = 1.0 / (1.0 - m * (x.sn * x.sn) * (d.sn * d.sn));
    return {(x.sn * d.cn * d.dn + d.sn * x.cn * x.dn) * den,
            (x.cn * d.cn - x.sn * x.dn * d.sn * d.dn) * den,
            (x.dn * d.dn - m * x.sn * x.cn * d.sn * d.cn) * den};
}
KS_DEV JS js_step_constants(double delta, double m) {
    JS x = js_at(fabs(delta), m);
    return {::copysign(x.sn, delta), x.cn, x.dn};
}
template <class B> KS_DEV JST<B> js_as(const JS& x) { return {B(x.sn), B(x.cn), B(x.dn)}; }

template <class R>
struct QCT {   // quadrature constants (quadrature.jl:42-97)
    R sp, sm, Qt, Qtm, Qph, Qphm, C, nq, sqrt_eps, gc, mpp;
    bool scattering;
};
using QC = QCT<double>;
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
template <class B> KS_DEV QCT<B> qc_as(const QC& q) {
    return {B(q.sp), B(q.sm), B(q.Qt), B(q.Qtm), B(q.Qph), B(q.Qphm), B(q.C), B(q.nq), B(q.sqrt_eps), B(q.gc), B(q.mpp), q.scattering};
}
template <class B> KS_DEV SpinT<B> spin_as(const Spin& S) { return {B(S.a), B(S.rp), B(S.rm)}; }
// the pixel constants in the bulk scalar (the flags and root case unchanged)
template <class B> KS_DEV PixCT<B> pix_as(const PixC& p) {
    PixCT<B> b;
    b.alpha = B(p.alpha); b.beta = B(p.beta); b.lam = B(p.lam); b.eta = B(p.eta); b.rcase = p.rcase; b.nreal = p.nreal;
    b.k = B(p.k); b.c = B(p.c); b.fo = B(p.fo); b.tau_total = B(p.tau_total); b.tau_h = B(p.tau_h);
    b.r1 = B(p.r1); b.r3 = B(p.r3); b.r4 = B(p.r4); b.n = B(p.n); b.n0 = B(p.n0); b.n1 = B(p.n1); b.d0 = B(p.d0); b.d1 = B(p.d1);
    b.a2 = B(p.a2); b.b1 = B(p.b1); b.g0 = B(p.g0); b.psi0 = B(p.psi0);
    b.Pp = B(p.Pp); b.Pm = B(p.Pm); b.Qp = B(p.Qp); b.Qm = B(p.Qm); b.It0 = B(p.It0); b.Iph0 = B(p.Iph0);
    b.up = B(p.up); b.um = B(p.um); b.f = B(p.f); b.mth = B(p.mth); b.off = B(p.off); b.sgn = B(p.sgn); b.npol = B(p.npol);
    b.Xth0 = B(p.Xth0); b.Gt0 = B(p.Gt0); b.Gphi0 = B(p.Gphi0); b.vortical = p.vortical; b.mu = B(p.mu); b.scale = B(p.scale); b.K2 = B(p.K2);
    return b;
}

// (rd, thd, cthd: the FP64 position the samples store and the visibility rule takes; r, th, cth the same in B)
template <class B> struct RPT { B r, th, cth, s, sn, cn, gt, gph; bool nur, nuth; double rd, thd, cthd; };
using RP = RPT<double>;

template <class B>
KS_DEV void smooth_integrands(const PixCT<B>& p, const SpinT<B>& S, const QCT<B>& q, B r, B cth, B s, B sn, B cn, B dn, B& gt, B& gph) {
    const B a = S.a, lam = p.lam;
    B D = r * r - 2.0 * r + a * a;
    B s2 = 1.0 - cth * cth;
    B iD = 1.0 / D;
    B ft = (r * r + a * a) * (r * r + a * a - a * lam) * iD + a * (lam - a * s2);
    B fph = a * (2.0 * r - a * lam) * iD;
    B wp = s * S.rp / (r * (r - S.rp)) / q.sp;
    B wm = s * S.rm / (r * (r - S.rm)) / q.sm;
    gt = ft - fabs(s) * (1.0 + 2.0 / r) - q.Qt * wp - q.Qtm * wm;
    B gpol = lam * q.C * (1.0 - q.gc * dn * (1.0 - q.mpp * cn * cn * 0.5)) / (1.0 - q.nq * sn * sn);
    gph = fph - q.Qph * wp - q.Qphm * wm + gpol;
}

// tau, fo: the Mino time of the point and of the radial turning point (FP64: they decide nu_r)
// (p64: the FP64 constants of the same pixel, for the polar angle of the FP32 bulk)
// The Jacobi states are FP64 for every B: FP32 addition-theorem steps drift by ~1e-7 per step, which over an anchor
// interval moved the emission radius by ~5e-5 of itself and the polar angle (near the axis, where acos amplifies the
// rounding of cos theta to ~sqrt(eps)) by more; the positions decide the splat weights. Their two steps per sample
// are a small part of the march; B carries the integrand rates and the Simpson increments.
template <class B>
KS_DEV RPT<B> ray_point(const PixC& p64, const PixCT<B>& p, const SpinT<B>& S, const QCT<B>& q, const JS& xr, const JS& xt, double tau,
                        double fo) {
    RPT<B> o;
    o.rd = radius_jac<double>(p64, xr.sn, xr.cn);
    polar_from_jac<double>(p64, xt.sn, xt.cn, xt.dn, o.thd, o.cthd, o.nuth);
    o.r = B(o.rd); o.th = B(o.thd); o.cth = B(o.cthd);
    o.nur = before_turning<double>(tau, fo);
    B R = r_potential<B>(S.a, p.lam, p.eta, o.r);
    const B w = o.r * o.r + S.a * S.a - S.a * p.lam;
    // R below 1e-14 w^2 is rounding (a sample on the radial turning point, where R = 0): s = 0 there
    o.s = (o.nur ? -1.0 : 1.0) * (R > 1e-14 * (w * w) ? sqrt(R) : B(0.0));
    const double idn = 1.0 / xt.dn;
    o.sn = B(xt.sn * idn / p64.scale);
    o.cn = B(xt.cn * idn);
    smooth_integrands<B>(p, S, q, o.r, o.cth, o.s, o.sn, o.cn, B(idn), o.gt, o.gph);
    return o;
}
KS_COLD JS radial_state(const PixC& p, double tau) { return js_at(p.c * (p.fo - tau), p.k); }
KS_COLD JS polar_state(const PixC& p, double tau) { return js_at(p.scale * (tau + p.off) / p.f, p.mu); }

KS_COLD RP ray_point_turning(const PixC& p, const Spin& S, const QC& q) {
    JS xt = polar_state(p, p.fo);
    RP o;
    o.r = p.r4;
    polar_from_jac(p, xt.sn, xt.cn, xt.dn, o.th, o.cth, o.nuth);
    o.rd = o.r; o.thd = o.th; o.cthd = o.cth;
    o.nur = true;
    o.s = 0.0;
    double idn = 1.0 / xt.dn;
    o.sn = xt.sn * idn / p.scale;
    o.cn = xt.cn * idn;
    smooth_integrands<double>(p, S, q, o.r, o.cth, o.s, o.sn, o.cn, idn, o.gt, o.gph);
    return o;
}
template <class B> KS_DEV RPT<B> rp_as(const RP& o) {
    return {B(o.r), B(o.th), B(o.cth), B(o.s), B(o.sn), B(o.cn), B(o.gt), B(o.gph), o.nur, o.nuth, o.rd, o.thd, o.cthd};
}

template <class B>
KS_DEV B polar_closed_form(const PixCT<B>& p, const QCT<B>& q, double tau, B sn, B cn) {
    B X = (B(tau) + p.off) / p.f;
    B nw = floor((X + p.K2 * 0.5) / p.K2);
    B A = (atan(q.sqrt_eps * sn / cn) + nw * PI) / q.sqrt_eps;
    B ph = atan(sn / cn) + nw * PI;
    B eps = q.sqrt_eps * q.sqrt_eps;
    return p.lam * q.C * p.f * q.gc * ((1.0 + q.mpp * eps / (2.0 * q.nq)) * A - q.mpp / (2.0 * q.nq) * ph);
}

// Pa, Pb: polar_closed_form at the interval's ends (each end's value serves both intervals that share it)
template <class B>
KS_DEV void interval_increment(const PixCT<B>& p, const SpinT<B>& S, const QCT<B>& q, const RPT<B>& pa, const RPT<B>& pm, const RPT<B>& pb,
                               double ta, double tb, B Pa, B Pb, B& It, B& Iph) {
    B h = B(tb - ta);
    B ra = pa.r, rb = pb.r;
    B dFr = (rb - ra) + 2.0 * log(rb / ra);
    B dFp = log((rb - S.rp) * ra / ((ra - S.rp) * rb)) / q.sp;
    B dFm = log((rb - S.rm) * ra / ((ra - S.rm) * rb)) / q.sm;
    It = h * (1.0 / 6.0) * (pa.gt + 4.0 * pm.gt + pb.gt) + fabs(dFr) + q.Qt * dFp + q.Qtm * dFm;
    Iph = h * (1.0 / 6.0) * (pa.gph + 4.0 * pm.gph + pb.gph) + q.Qph * dFp + q.Qphm * dFm + Pb - Pa;
}

// the interval holding the radial turning point of a scattering ray, split there (quadrature.jl:255-270); the
// turning point and the two midpoints come from the closed forms (FP64), the increments in the bulk scalar
template <class B>
KS_COLD void turning_increment(const PixC& p, const Spin& S, const QC& q, const PixCT<B>& pb_, const SpinT<B>& Sb, const QCT<B>& qb,
                               const RPT<B>& pa, const RPT<B>& pb, double ta, double tb, B Pa, B Pb, double& t, double& ph) {
    const RPT<B> pt = rp_as<B>(ray_point_turning(p, S, q));
    const B Pt = polar_closed_form<B>(pb_, qb, p.fo, pt.sn, pt.cn);
    double t1 = 0.5 * (ta + p.fo), t2 = 0.5 * (p.fo + tb);
    RPT<B> p1 = ray_point<B>(p, pb_, Sb, qb, radial_state(p, t1), polar_state(p, t1), t1, p.fo);
    RPT<B> p2 = ray_point<B>(p, pb_, Sb, qb, radial_state(p, t2), polar_state(p, t2), t2, p.fo);
    B It1, Ip1, It2, Ip2;
    interval_increment<B>(pb_, Sb, qb, pa, p1, pt, ta, p.fo, Pa, Pt, It1, Ip1);
    interval_increment<B>(pb_, Sb, qb, pt, p2, pb, p.fo, tb, Pt, Pb, It2, Ip2);
    t += val(It1) + val(It2);
    ph += val(Ip1) + val(Ip2);
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
// B: the bulk scalar (above). The running t~ and phi are FP64 sums of the B-scalar increments between the FP64
// anchors; the samples handed to f are FP64 (converted from B).
// (pb_, Sb, qb: the pixel, spin and quadrature constants in B; for B = double the FP64 ones themselves. Krang's
// visibility rule takes the polar angle in FP64: its cos^2 theta_o - cos^2 theta cancels near the threshold.)
template <class B, class F>
KS_DEV void march_range_b(const PixC& p, const Spin& S, const QC& q, const VisC& vis, const PixCT<B>& pb_, const SpinT<B>& Sb, const QCT<B>& qb,
                          int N, int M, int kbeg, int kend, F& f, double& resid_t, double& resid_ph) {
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
    RPT<B> pa;
    for (;;) {
        const double ta = (double)A * dtau;
        xr = radial_state(p, ta);
        xt = polar_state(p, ta);
        pa = ray_point<B>(p, pb_, Sb, qb, xr, xt, ta, p.fo);
        if (A <= 1 || ((ta <= p.tau_h) && krang_visible(vis, pa.cthd))) break;
        A = A - Man > 1 ? A - Man : 1;
    }
    double t, ph;
    t_phi_direct(p, S, (double)A * dtau, t, ph);
    B Pa = polar_closed_form<B>(pb_, qb, (double)A * dtau, pa.sn, pa.cn);
    if (A == kbeg) {
        const double ta = (double)A * dtau;
        Samp s{t, pa.rd, pa.thd, ph, pa.nur, pa.nuth, (ta <= p.tau_h) && krang_visible(vis, pa.cthd)};
        f(A, s, dtau);
    }
    const int klast = kend < N ? kend + 1 : N;
    for (int k = A + 1; k <= klast; ++k) {
        const double ta = (double)(k - 1) * dtau, tb = (double)k * dtau;
        const bool anchor = ((k - 1) % Man) == 0;
        if (k > kend && !anchor) break;            // the step past kend only serves the next anchor's residual
        xr = js_step<double>(xr, Dr, p.k);
        xt = js_step<double>(xt, Dt, p.mu);
        RPT<B> pmid = ray_point<B>(p, pb_, Sb, qb, xr, xt, ta + h, p.fo);
        if (anchor) { xr = radial_state(p, tb); xt = polar_state(p, tb); }
        else { xr = js_step<double>(xr, Dr, p.k); xt = js_step<double>(xt, Dt, p.mu); }
        RPT<B> pb = ray_point<B>(p, pb_, Sb, qb, xr, xt, tb, p.fo);
        const B Pb = polar_closed_form<B>(pb_, qb, tb, pb.sn, pb.cn);
        if (q.scattering && p.fo - ta > tol && tb - p.fo > tol) {
            turning_increment<B>(p, S, q, pb_, Sb, qb, pa, pb, ta, tb, Pa, Pb, t, ph);
        } else {
            B It, Ip;
            interval_increment<B>(pb_, Sb, qb, pa, pmid, pb, ta, tb, Pa, Pb, It, Ip);
            t += val(It);
            ph += val(Ip);
        }
        const bool okb = (tb <= p.tau_h) && krang_visible(vis, pb.cthd);
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
            Samp s{t, pb.rd, pb.thd, ph, pb.nur, pb.nuth, okb};
            f(k, s, dtau);
        }
        pa = pb;
        Pa = Pb;
    }
}
template <class B, class F>
KS_DEV void march_range(const PixC& p, const Spin& S, double thetao, int N, int M, int kbeg, int kend, F& f, double& resid_t, double& resid_ph) {
    const QC q = make_qc(p, S);
    const VisC vis = vis_constants(p, S.a, thetao);
    if constexpr (std::is_same<B, double>::value) {
        march_range_b<double, F>(p, S, q, vis, p, S, q, N, M, kbeg, kend, f, resid_t, resid_ph);
    } else {
        const PixCT<B> pb_ = pix_as<B>(p);
        march_range_b<B, F>(p, S, q, vis, pb_, spin_as<B>(S), qc_as<B>(q), N, M, kbeg, kend, f, resid_t, resid_ph);
    }
}
template <class F>
KS_DEV void march_range(const PixC& p, const Spin& S, double thetao, int N, int M, int kbeg, int kend, F& f, double& resid_t, double& resid_ph) {
    march_range<double, F>(p, S, thetao, N, M, kbeg, kend, f, resid_t, resid_ph);
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
