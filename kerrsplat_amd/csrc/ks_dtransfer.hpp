// Device polarized synchrotron transfer: per-splat constants, the splat weight and bounding-sphere
// cull, the fluid frame and Walker-Penrose screen angle, thermal / power-law / kappa coefficients,
// the polarization cap and nu-hat invariants, and the exact constant-coefficient 4x4 Stokes step.
// Everything is a template on its scalar: double (the FP64 path), f32 (the FP32 path) or duals of
// either (the adjoint); see ks_ddual.hpp.
//
// Reference: src/Transfer/{coefficients,kappa,step,frames,transport}.jl and
// src/Splats/{Splats,polarized,populations,trajectories}.jl. The synchrotron fits are the published
// ones (Dexter 2016, Shcherbakov 2008, Pandya+ 2016, Marszewski+ 2021, Jones & O'Dell 1977) and are
// written out as such; everything that depends only on a splat (Bessel ratios at 1/Theta_e, Gamma
// functions, the kappa 2F1 factor, the quaternion rotation, field direction, boost) is computed
// once per splat on the host instead of at every sample.
#pragma once
#include "ks_dgeo.hpp"
#include "ks_ddual.hpp"

namespace ksd {

constexpr double ME = 9.1093826e-28, CL = 2.99792458e10, EE = 4.80320680e-10, HPL = 6.6260693e-27;
constexpr double NU0 = 1e11;
constexpr double JY = 1e-23;                  // erg s^-1 cm^-2 Hz^-1
constexpr double PLANCK0 = 2.0 * HPL * NU0 * NU0 * NU0 / (CL * CL);
constexpr double WEIGHT_CUTOFF = 1e-6;
// an exponent above GEXP_SKIP leaves the Gaussian weight exp(-gexp / 2) below the cutoff in any precision
// (exp(-13.82) = 9.96e-7: a margin of 0.4%, far above FP32 rounding), so the exponential is not taken there
constexpr double GEXP_SKIP = 27.64;

// MODEL_COMPOSITE: concatenated thermal, power-law and kappa splats (CompositeModel, transport.jl:298-316),
// each splat's population in SplatCT::kind
enum ModelKind : int { MODEL_THIN = 0, MODEL_THERMAL = 1, MODEL_POWERLAW = 2, MODEL_KAPPA = 3, MODEL_KNOT = 4, MODEL_COMPOSITE = 5 };

// The kappa fits' factors that depend on the splat alone (kappa, w, the 2F1 factor, the Gamma values, |B|), built
// once per splat by kappa_consts; kappa_coefs evaluates the fits at one sample from them. Indices into KappaK::v:
enum KappaField : int {
    KK_NUW1,                       // (w kappa)^2 nu_c: nu_w = KK_NUW1 sin(theta), X = nu / nu_w
    KK_JPRE, KK_APRE,              // e^2 nu_c / c (emissivity prefactor / (n_e sin)), e / B (absorptivity's, x sin / n_e)
    KK_NLOWI, KK_RI, KK_EI, KK_MI, KK_XI,       // N_low; (N_low / N_high)^x; x (3k - 4) / 6; -1 / x; x   (j_I; x_V = x_I)
    KK_NLOWQ, KK_RQ, KK_EQ, KK_MQ,              // j_Q
    KK_CV1, KK_CV2,                // N_low,V = CV1 (sin^-2.4 - 1)^0.48 X^-0.35, N_high,V = CV2 (sin^-2.5 - 1)^0.44 / sqrt(X)
    KK_LOWA, KK_RA, KK_EA, KK_MA,               // alpha_I (x (3k - 1) / 6)
    KK_NLOWQA, KK_RQA, KK_EQA, KK_MQA,          // alpha_Q
    KK_CVA1, KK_CVA2, KK_XVA, KK_EVA, KK_MVA,   // alpha_V: N_low = CVA1 (sin^-2.28 - 1)^0.446 X^-0.35, N_high = CVA2 sqrt(sin^-2.05 - 1) / sqrt(X)
    KK_FQ0, KK_FQ1, KK_FV0, KK_FV1,             // rotativity fits' w-dependent weights at the fit below kappa and the next
    KK_RHOQ, KK_RHOV,              // e^2 nu_c^2 / (m c); 2 e^2 nu_c K0/K2 / (m c)
    KK_IFIT,                       // the rotativity fit below kappa (1..3)
    KK_N
};
template <class T> struct KappaK { T v[KK_N]; };

// Per-splat constants, built on the host by ks_model_create in both precisions (SplatC, SplatCf);
// a kernel reads them with wave-uniform addresses.
template <class S> struct SplatCT {
    S x, y, z;            // centre at t0 (rows 1-3)
    S R[9];               // quaternion rotation, row-major R[a][b]
    S is[3];              // exp(-s_a)
    S t0, iw;             // t0 and exp(-logw)
    S omega;              // pattern angular velocity
    S rsup2;              // support sphere radius^2 = -2 ln(cutoff) exp(2 max s)
    S lnA;                // thin: logA; others: logne
    S Bmag, bh[3];        // field magnitude and unit direction (ZAMO axes r, phi, -theta)
    S u[3];               // fluid spatial 4-velocity (ZAMO axes)
    S p1, p2, p3;         // thermal: Theta_e, K0/K2, K1/K2 at 1/Theta_e; power law: p, gamma_min, gamma_max;
                          // kappa: kappa, w, hyp
    S g[6];               // power law: Gamma((3p-1)/12) Gamma((3p+19)/12), Gamma((3p+2)/12) Gamma((3p+22)/12);
                          // kappa: Gamma(k-4/3), Gamma(k-2), Gamma(k/4-1/3), Gamma(k/4+4/3), Gamma(2+k/2), K0/(K2+1e-40) at 1/w
    S kc[3], kr;          // splats whose centre follows knots: a sphere holding every knot (centre, radius; kr = 0 without)
    KappaK<S> kk;         // kappa: the fits' per-splat factors (kappa_consts)
    int knot0, nk, kt0;   // splats whose centre follows knots (nk > 0): the first of its nk knots (xyz triples) and of their
                          // times in the model's arrays
    int kind;             // the splat's population (MODEL_THERMAL, MODEL_POWERLAW, MODEL_KAPPA, ...)
};
using SplatC = SplatCT<double>;
using SplatCf = SplatCT<f32>;

// ---- geometry ---------------------------------------------------------------------------------------------
template <class Q>
KS_DEV void quasi_cartesian_ks(Q a, Q rp, Q rm, Q r, Q th, Q ph, Q& x, Q& y, Q& z) {
    Q temp = sqrt(1.0 - a * a);
    Q phks = ph + a / (2.0 * temp) * log(fabs((r - rp + EPS) / (r - rm + EPS))) - atan2(a, r);
    Q st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(phks, &sp, &cp);
    x = r * st * cp; y = r * st * sp; z = r * ct;
}

// centre of a splat at time t under the pattern rotation and the squared distance (for the cull)
template <class S>
KS_DEV bool outside_support(const SplatCT<S>& s, S t, S x, S y, S z, S& sp, S& cp) {
    S ph = s.omega * (t - s.t0);
    sincos(ph, &sp, &cp);
    S cx = s.x * cp - s.y * sp, cy = s.x * sp + s.y * cp;
    S dx = x - cx, dy = y - cy, dz = z - s.z;
    return dx * dx + dy * dy + dz * dz > s.rsup2;
}
// Whether a splat can reach a region of samples (the per-tile cull): the region is a sphere (centre p,
// radius R) around its sample positions, with emission times in [tmin, tmax]. Over those times the
// splat's centre sweeps an arc of the circle about the spin axis; every point of an arc of half-angle h
// is within 2 rho sin(min(h, pi) / 2) of the arc's midpoint. False only when no sample of the region
// can be inside the support sphere (outside_support), with a margin for rounding.
// A centre that follows knots (knots, ktimes: the model's arrays, nullptr to use the sphere holding every knot)
// moves over [tmin, tmax] along the polyline through its positions at tmin and tmax and the knots between (linear
// between knots, held outside them): it stays inside the bounding box of those points, bounded by the sphere
// around the box's centre through the farthest of them.
template <class S>
KS_DEV bool splat_may_reach(const SplatCT<S>& c, double px, double py, double pz, double R, double tmin, double tmax,
                            const S* knots = nullptr, const S* ktimes = nullptr) {
    if (c.nk > 0) {
        double cx = val(c.kc[0]), cy = val(c.kc[1]), cz = val(c.kc[2]), cr = val(c.kr);
        if (knots && ktimes) {
            const S* K = knots + 3 * (int64_t)c.knot0;
            const S* kt = ktimes + c.kt0;
            const int nk = c.nk;
            double pt[2][3];
            for (int e = 0; e < 2; ++e) {
                const double t = e == 0 ? tmin : tmax;
                if (!(t > val(kt[0]))) { for (int d = 0; d < 3; ++d) pt[e][d] = val(K[d]); }
                else if (t >= val(kt[nk - 1])) { for (int d = 0; d < 3; ++d) pt[e][d] = val(K[3 * (nk - 1) + d]); }
                else {
                    int kk = 0;
                    while (kk < nk - 2 && val(kt[kk + 1]) <= t) ++kk;
                    const double w = (t - val(kt[kk])) / (val(kt[kk + 1]) - val(kt[kk]));
                    for (int d = 0; d < 3; ++d) pt[e][d] = (1.0 - w) * val(K[3 * kk + d]) + w * val(K[3 * (kk + 1) + d]);
                }
            }
            double lo[3], hi[3];
            for (int d = 0; d < 3; ++d) { lo[d] = fmin(pt[0][d], pt[1][d]); hi[d] = fmax(pt[0][d], pt[1][d]); }
            for (int k = 0; k < nk; ++k) {
                const double tk = val(kt[k]);
                if (tk > tmin && tk < tmax)
                    for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], val(K[3 * k + d])); hi[d] = fmax(hi[d], val(K[3 * k + d])); }
            }
            cx = 0.5 * (lo[0] + hi[0]); cy = 0.5 * (lo[1] + hi[1]); cz = 0.5 * (lo[2] + hi[2]);
            cr = 0.5 * sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) + (hi[2] - lo[2]) * (hi[2] - lo[2]));
        }
        const double dx = px - cx, dy = py - cy, dz = pz - cz;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double reach = R + sqrt(val(c.rsup2)) + cr;
        // margin: the per-sample test in FP32 rounds the emission time (which moves the interpolated centre by
        // |dc/dt| 6e-8 |t|) and the positions; 1e-4 relative covers them
        return d <= reach + 1e-4 * (reach + d + cr) + 1e-6;
    }
    const double x = val(c.x), y = val(c.y), z = val(c.z), w = val(c.omega);
    const double tm = 0.5 * (tmin + tmax), h = 0.5 * fabs(w) * (tmax - tmin);
    double sp, cp;
    sincos(w * (tm - val(c.t0)), &sp, &cp);
    const double cx = x * cp - y * sp, cy = x * sp + y * cp;
    const double arc = 2.0 * sqrt(x * x + y * y) * sin(0.5 * dmin(h, PI));
    const double dx = px - cx, dy = py - cy, dz = pz - z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double rho = sqrt(x * x + y * y);
    const double reach = R + sqrt(val(c.rsup2)) + arc;
    // margin: the per-sample test in FP32 rounds the emission time (|t| 6e-8, turned by omega into a phase
    // error) and the positions; 1e-4 relative covers the rest
    const double margin = 1e-4 * (reach + d) + 4e-7 * rho * fabs(w) * dmax(fabs(tmin), fabs(tmax)) + 1e-6;
    return d <= reach + margin;
}

// the Gaussian exponent q^2 + tau^2 given the rotation phase (Splats.jl:58-93)
template <class S>
KS_DEV S gauss_exponent(const SplatCT<S>& s, S t, S x, S y, S z, S sp, S cp) {
    S cx = s.x * cp - s.y * sp, cy = s.x * sp + s.y * cp;
    S d0 = x - cx, d1 = y - cy, d2 = z - s.z;
    S e0 = d0 * cp + d1 * sp, e1 = -d0 * sp + d1 * cp, e2 = d2;
    S u0 = s.R[0] * e0 + s.R[3] * e1 + s.R[6] * e2;
    S u1 = s.R[1] * e0 + s.R[4] * e1 + s.R[7] * e2;
    S u2 = s.R[2] * e0 + s.R[5] * e1 + s.R[8] * e2;
    S q0 = u0 * s.is[0], q1 = u1 * s.is[1], q2 = u2 * s.is[2];
    S tt = (t - s.t0) * s.iw;
    return q0 * q0 + q1 * q1 + q2 * q2 + tt * tt;
}

// The centre of a splat that follows knots at emission time t (trajectories.jl:93-111): linear interpolation
// between the knots around t, held at the first and last knot outside their times.
template <class S>
KS_DEV void knot_centre(const SplatCT<S>& sc, const S* __restrict__ knots, const S* __restrict__ ktimes, S t, S& cx, S& cy, S& cz) {
    const S* K = knots + 3 * (int64_t)sc.knot0;
    const S* kt = ktimes + sc.kt0;
    const int nk = sc.nk;
    if (t <= kt[0]) { cx = K[0]; cy = K[1]; cz = K[2]; }
    else if (t >= kt[nk - 1]) { cx = K[3 * (nk - 1)]; cy = K[3 * (nk - 1) + 1]; cz = K[3 * (nk - 1) + 2]; }
    else {
        int kk = 0;
        while (kk < nk - 2 && kt[kk + 1] <= t) ++kk;
        S w = (t - kt[kk]) / (kt[kk + 1] - kt[kk]);
        cx = (1.0 - w) * K[3 * kk] + w * K[3 * (kk + 1)];
        cy = (1.0 - w) * K[3 * kk + 1] + w * K[3 * (kk + 1) + 1];
        cz = (1.0 - w) * K[3 * kk + 2] + w * K[3 * (kk + 1) + 2];
    }
}

// ---- metric, ZAMO tetrad, photon momentum ----------------------------------------------------------------------
template <class T> struct GeoT {   // everything of a sample the splats share
    T r, th, st, ct, S, Dl, Xi, lapse, omega, sgrr, sgthth, sgpp;
    T pu[4];        // contravariant BL photon momentum
    T pz[4];        // ZAMO components (t, r, phi, -theta)
};
using Geo = GeoT<double>;
template <class Q>
KS_DEV void make_geo(Q a, Q r, Q th, Q eta, Q lam, bool nur, bool nuth, GeoT<Q>& g) {
    g.r = r; g.th = th;
    sincos(th, &g.st, &g.ct);
    g.S = r * r + a * a * g.ct * g.ct;
    g.Dl = r * r - 2.0 * r + a * a;
    g.Xi = (r * r + a * a) * (r * r + a * a) - a * a * g.Dl * g.st * g.st;
    g.lapse = sqrt(g.Dl * g.S / g.Xi);
    g.omega = 2.0 * a * r / g.Xi;
    g.sgrr = sqrt(g.S / g.Dl);
    g.sgthth = sqrt(g.S);
    g.sgpp = g.st * sqrt(g.Xi / g.S);
    Q w = r * r + a * a - a * lam;
    Q R = w * w - g.Dl * (eta + (lam - a) * (lam - a));
    Q Th = eta + a * a * g.ct * g.ct - lam * lam * g.ct * g.ct / (g.st * g.st);
    // R below 1e-14 w^2 is rounding (a sample on the radial turning point): its square root would be noise
    Q pr = (nur ? 1.0 : -1.0) * (R > 1e-14 * (w * w) ? sqrt(R) : Q(0.0)) / g.Dl;
    Q pth = (nuth ? 1.0 : -1.0) * (Th > 0.0 ? sqrt(Th) : Q(0.0));
    // p^mu = g^{mu nu} p_nu with p_t = -1, p_phi = lam
    Q DS = g.Dl * g.S;
    Q gtt = -g.Xi / DS, gtp = -2.0 * a * r / DS, gpp = (g.Dl - a * a * g.st * g.st) / (DS * g.st * g.st);
    g.pu[0] = -gtt + gtp * lam;
    g.pu[1] = g.Dl / g.S * pr;
    g.pu[2] = pth / g.S;
    g.pu[3] = -gtp + gpp * lam;
    g.pz[0] = g.lapse * g.pu[0];
    g.pz[1] = g.sgrr * g.pu[1];
    g.pz[2] = g.sgpp * (g.pu[3] - g.omega * g.pu[0]);
    g.pz[3] = -g.sgthth * g.pu[2];
}
// The fields of a sample's geometry that the frames use (local_frame, frame_jacobian), stored field-major at
// w[f * st] (f < 16): pz, pu, lapse, sgrr, sgthth, sgpp, omega, r, st, ct. geo_load leaves th, S, Dl and Xi unset.
template <class T> KS_DEV void geo_store(const GeoT<T>& g, T* w, int64_t st) {
    for (int i = 0; i < 4; ++i) { w[i * st] = g.pz[i]; w[(4 + i) * st] = g.pu[i]; }
    w[8 * st] = g.lapse; w[9 * st] = g.sgrr; w[10 * st] = g.sgthth; w[11 * st] = g.sgpp; w[12 * st] = g.omega;
    w[13 * st] = g.r; w[14 * st] = g.st; w[15 * st] = g.ct;
}
template <class T> KS_DEV void geo_load(const T* o, int64_t st, GeoT<T>& g) {
    for (int i = 0; i < 4; ++i) { g.pz[i] = o[i * st]; g.pu[i] = o[(4 + i) * st]; }
    g.lapse = o[8 * st]; g.sgrr = o[9 * st]; g.sgthth = o[10 * st]; g.sgpp = o[11 * st]; g.omega = o[12 * st];
    g.r = o[13 * st]; g.st = o[14 * st]; g.ct = o[15 * st];
    g.th = T(0.0); g.S = T(0.0); g.Dl = T(0.0); g.Xi = T(0.0);
}
// the FP64 geometry of a sample in another scalar (built once per sample, shared by the splats)
template <class T> KS_DEV GeoT<T> geo_as(const Geo& g) {
    GeoT<T> o;
    o.r = T(g.r); o.th = T(g.th); o.st = T(g.st); o.ct = T(g.ct); o.S = T(g.S); o.Dl = T(g.Dl); o.Xi = T(g.Xi);
    o.lapse = T(g.lapse); o.omega = T(g.omega); o.sgrr = T(g.sgrr); o.sgthth = T(g.sgthth); o.sgpp = T(g.sgpp);
    for (int i = 0; i < 4; ++i) { o.pu[i] = T(g.pu[i]); o.pz[i] = T(g.pz[i]); }
    return o;
}

template <class T> struct FrameT { T g, cosB, chi; };
using Frame = FrameT<double>;

// frames.jl:108-128 for one splat: its boost built from the fluid velocity u, the field direction bh
// (unit); T = the scalar S of the geometry, or duals over it in u and bh for the adjoint
template <class T, class S>
KS_DEV FrameT<T> local_frame(const GeoT<S>& G, S a, S alpha, S beta, S thetao_sin, const T u[3], const T bh[3]) {
    T u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    T gam = sqrt(1.0 + u2);
    T q = 1.0 / (gam + 1.0);
    // p_f = Lambda p_z
    T ud = u[0] * G.pz[1] + u[1] * G.pz[2] + u[2] * G.pz[3];
    T pf0 = gam * G.pz[0] - ud;
    T c1 = ud * q - G.pz[0];
    T pf1 = G.pz[1] + u[0] * c1, pf2 = G.pz[2] + u[1] * c1, pf3 = G.pz[3] + u[2] * c1;
    FrameT<T> fr;
    fr.g = 1.0 / pf0;
    T k0 = pf1 * fr.g, k1 = pf2 * fr.g, k2 = pf3 * fr.g;
    fr.cosB = k0 * bh[0] + k1 * bh[1] + k2 * bh[2];
    T f0 = k1 * bh[2] - k2 * bh[1], f1 = k2 * bh[0] - k0 * bh[2], f2 = k0 * bh[1] - k1 * bh[0];
    T fn = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
    if (!(fn > 0.0)) { fr.chi = T(0.0); return fr; }
    T ifn = 1.0 / fn;
    f0 = f0 * ifn; f1 = f1 * ifn; f2 = f2 * ifn;
    // inverse boost of (0, f): time part = u.f, spatial f + u (u.f) q
    T uf = u[0] * f0 + u[1] * f1 + u[2] * f2;
    T ufq = uf * q;
    T fz1 = f0 + u[0] * ufq, fz2 = f1 + u[1] * ufq, fz3 = f2 + u[2] * ufq;
    // ZAMO -> BL contravariant
    T ft = uf * (1.0 / G.lapse);
    T fr_ = fz1 * (1.0 / G.sgrr);
    T fth = fz3 * (-1.0 / G.sgthth);
    T fph = fz2 * (1.0 / G.sgpp) + G.omega * ft;
    const S* p = G.pu;
    T A = p[0] * fr_ - p[1] * ft + (a * G.st * G.st) * (p[1] * fph - p[3] * fr_);
    T Bw = ((G.r * G.r + a * a) * (p[3] * fth - p[2] * fph) - a * (p[0] * fth - p[2] * ft)) * G.st;
    T kk1 = A * G.r - Bw * (a * G.ct), kk2 = -(A * (a * G.ct) + Bw * G.r);
    S mu = -(alpha + a * thetao_sin);
    T ea = beta * kk2 - mu * kk1, eb = beta * kk1 + mu * kk2;
    fr.chi = atan2(-ea, eb);
    return fr;
}

// sin and cos of the pitch angle acos(cb), cb in [-1, 1] (the reference's safe_acos, then its sin and cos):
// cos = cb and sin = sqrt((1 - cb)(1 + cb)), whose product form stays exact to an ulp near |cb| = 1
template <class T> KS_DEV void pitch_sincos(const T& cb, T& sth, T& cth) {
    cth = cb;
    sth = sqrt((1.0 - cb) * (1.0 + cb));
}

// ---- coefficients ------------------------------------------------------------------------------------------------
// The power-law and kappa fits diverge as the pitch-angle sine goes to zero (a photon along the field). A sine that
// rounds to zero (FP32 does so below about 3.5e-4 rad, FP64 below 1.5e-8 rad) is taken as half the smallest nonzero
// sine the scalar resolves next to |cos| = 1, sqrt(eps) / 2, as a constant (zero partials).
KS_DEV double sin_floor(double) { return 7.45e-9; }
KS_DEV double sin_floor(f32) { return 1.73e-4; }
template <int N, class S> KS_DEV double sin_floor(const DN<N, S>&) { return sin_floor(S(0.0)); }
// sin^-p - 1 of the pitch angle: near sin = 1 from the cosine, expm1(-(p / 2) log1p(-cos^2)), without the
// cancellation of the difference (which rounds to zero, and leaves the partials of x^e at zero unbounded, when the
// sine rounds to one); directly elsewhere
template <class T> KS_DEV T sin_pow_m1(T sth, T cth, double p) {
    if (cth * cth < 0.5) return expm1((-0.5 * p) * log1p(-(cth * cth)));
    return pow(sth, -p) - 1.0;
}
// x^e (0 < e < 1) of x >= 0, zero with zero partials at x = 0, where the derivative is unbounded (the reference's
// safe_sqrt makes the same choice for the square root, frames.jl:83-97)
template <class T> KS_DEV T pow_nonneg(T x, double e) { return x > 0.0 ? pow(x, e) : T(0.0); }

template <class T> struct CoefT { T jI, jQ, jV, aI, aQ, aV, rQ, rV; };
using Coef = CoefT<double>;

template <class T> KS_DEV T planck(T nu, T Th) {
    T x = HPL * nu / (ME * CL * CL * Th);
    T nh = nu * (1.0 / NU0);
    T pref = PLANCK0 * nh * nh * nh;
    if (x < 2e-3) return pref / (x / 24.0 * (24.0 + x * (12.0 + x * (4.0 + x))));
    return pref / expm1(x);
}

// thermal (coefficients.jl:101-151) with the splat's Bessel ratios k0/k2, k1/k2 (k2 > 0 in practice);
// ne enters linearly and stays a plain number in the adjoint
template <class T, class NE>
KS_DEV CoefT<T> thermal(NE ne, T Th, T k0k2, T k1k2, T B, T nu, T sth, T cth) {
    CoefT<T> c;
    T w = 2.0 * PI * nu;
    T w0w = EE * B / (ME * CL) / w;
    T wp2w2 = (4.0 * PI * ne * EE * EE / ME) / (w * w);
    if (!(sth > 0.0)) {
        c.jI = c.jQ = c.jV = c.aI = c.aQ = c.aV = c.rQ = T(0.0);
        c.rV = w / CL * (wp2w2 * w0w * k0k2 * cth);
        return c;
    }
    T x = Th * sqrt(1.4142135623730951 * sth * (1e3 * w0w));
    T lx = log(x);
    T extraterm = (0.011 * exp(-x / 47.2) - 0.7937005259840998 / 67.44733739010397 * PI * 1e4 * exp(-8.0 / 3.0 * log(x + 1e-16))) *
                  (0.5 + 0.5 * tanh((lx - 4.787491742782046) / 0.1));
    T jffunc = 2.011 * exp(-exp(1.035 * lx) / 4.7) - cos(x * 0.5) * exp(-exp(1.2 * lx) / 2.73) - 0.011 * exp(-x / 47.2) + extraterm;
    T eps11m22 = jffunc * wp2w2 * w0w * w0w * (k1k2 + 6.0 * Th) * sth * sth;
    c.rQ = w / (2.0 * CL) * eps11m22;
    T ff = k0k2 * (1.0 - 0.11 * log1p(0.035 * x));
    c.rV = w / CL * (wp2w2 * w0w * ff * cth);
    T nus = 3.0 * EE * B * sth / (4.0 * PI * ME * CL) * Th * Th + 1.0;
    T xs = nu / nus;
    T x13 = cbrt(xs), xm13 = 1.0 / x13, xm23 = xm13 * xm13;
    T e = exp(-1.8899 * x13);
    T pref = ne * EE * EE * nu / (2.0 * 1.7320508075688772 * CL * Th * Th);
    c.jI = pref * 2.5651 * (1.0 + 1.92 * xm13 + 0.9977 * xm23) * e;
    c.jQ = pref * 2.5651 * (1.0 + 0.93193 * xm13 + 0.499873 * xm23) * e;
    c.jV = 2.0 * ne * EE * EE * nu * cth / sth / (3.0 * 1.7320508075688772 * CL * Th * Th * Th) *
           (1.81384 / xs + 3.42319 * xm23 + 0.0292545 / sqrt(xs) + 2.03773 * xm13) * e;
    T Bn = planck(nu, Th);
    T ib = Bn > 0.0 ? 1.0 / Bn : T(0.0);
    c.aI = c.jI * ib; c.aQ = c.jQ * ib; c.aV = c.jV * ib;
    return c;
}

// power law (coefficients.jl:208-256); gPL = {Gamma((3p-1)/12) Gamma((3p+19)/12), Gamma((3p+2)/12) Gamma((3p+22)/12)}
template <class T>
KS_DEV CoefT<T> powerlaw(T ne, T p, T gmin, T gmax, const T* gPL, T B, T nu, T sth, T cth) {
    CoefT<T> c;
    if (!(sth > 0.0)) sth = T(sin_floor(sth));
    T nuc = EE * B / (2.0 * PI * ME * CL);
    T gspan = pow(gmin, 1.0 - p) - pow(gmax, 1.0 - p);
    T pref = ne * EE * EE * nuc / CL;
    T X = nu / (nuc * sth);
    c.jI = pref * (pow(3.0, p * 0.5) * (p - 1.0) * sth) / (2.0 * (p + 1.0) * gspan) * gPL[0] * pow(X, -(p - 1.0) * 0.5);
    c.jQ = (p + 1.0) / (p + 7.0 / 3.0) * c.jI;
    c.jV = (171.0 / 250.0) * pow(p, 0.49) * cth / sth * pow(nu / (3.0 * nuc * sth), -0.5) * c.jI;
    T apref = ne * EE * EE / (nu * ME * CL);
    c.aI = apref * pow(3.0, (p + 1.0) * 0.5) * (p - 1.0) / (4.0 * gspan) * gPL[1] * pow(X, -(p + 2.0) * 0.5);
    c.aQ = c.aI * pow((17.0 / 500.0) * p - 43.0 / 1250.0, 43.0 / 500.0);
    T term6 = pow_nonneg(3.1 * sin_pow_m1(sth, cth, 48.0 / 25.0), 64.0 / 125.0);
    const double sgn = cth > 0.0 ? 1.0 : (cth < 0.0 ? -1.0 : 0.0);
    c.aV = c.aI * pow(0.71 * p + 22.0 / 625.0, 197.0 / 500.0) * term6 * pow(X, -0.5) * sgn;
    T nuB = nuc * sth;
    T rperp = ne * EE * EE * (p - 1.0) / (ME * CL * nuB * gspan);
    T numin = gmin * gmin * nuB;
    T rb = nuB / nu;
    c.rQ = -rperp * rb * rb * rb * pow(gmin, 2.0 - p) * (1.0 - pow(numin / nu, p * 0.5 - 1.0)) / (p * 0.5 - 1.0);
    c.rV = 2.0 * (p + 2.0) / (p + 1.0) * rperp * rb * rb * pow(gmin, -(p + 1.0)) * log(gmin) * cth / sth;
    return c;
}

// kappa (kappa.jl:25-119), split into its per-splat factors and its per-sample evaluation; gK = {G(k-4/3),
// G(k-2), G(k/4-1/3), G(k/4+4/3), G(2+k/2), K0/(K2+1e-40) at 1/w}; B the field magnitude
template <class T>
KS_DEV void kappa_consts(T kap, T w, T hyp, const T* gK, T B, KappaK<T>& K) {
    const double small = 1e-40;
    T nuc = EE * B / (2.0 * PI * ME * CL);
    K.v[KK_NUW1] = (w * kap) * (w * kap) * nuc;
    K.v[KK_JPRE] = EE * EE * nuc / CL;
    K.v[KK_APRE] = EE / B;
    T g43 = gK[0], g2 = gK[1], gq = gK[2], gp = gK[3], gk2 = gK[4];
    const double c37 = 12.980246132766677;      // 3^(7/3)
    T NlowI = 4.0 * PI * g43 / (c37 * g2);
    T hi = 0.25 * pow(3.0, (kap - 1.0) * 0.5) * (kap - 2.0) * (kap - 1.0) * gq * gp;
    T xI = 3.0 * pow(kap, -1.5);
    K.v[KK_NLOWI] = NlowI;
    K.v[KK_RI] = pow(NlowI / (hi + small), xI);
    K.v[KK_EI] = xI * (3.0 * kap - 4.0) / 6.0;
    K.v[KK_MI] = -1.0 / xI;
    K.v[KK_XI] = xI;
    T NlowQ = -0.5 * NlowI;
    T NhighQ = -(0.64 + kap / 50.0) * hi + small;
    T xQ = 3.7 * pow(kap, -1.6);
    K.v[KK_NLOWQ] = NlowQ;
    K.v[KK_RQ] = pow(NlowQ / NhighQ, xQ);
    K.v[KK_EQ] = xQ * (3.0 * kap - 4.0) / 6.0;
    K.v[KK_MQ] = -1.0 / xQ;
    K.v[KK_CV1] = -0.5625 * (pow(kap, -0.528) / w) * NlowI;
    K.v[KK_CV2] = -0.765625 * (pow(kap, -0.44) / w) * hi;
    const double g53 = 0.9027452929509338;     // Gamma(5/3)
    T ck = (kap - 2.0) * (kap - 1.0) * kap;
    const double c16 = 1.2009369551760027;     // 3^(1/6)
    const double tpi2 = 39.47841760435743;     // (2 pi)^2
    T wk = w * kap;
    T lowA = c16 * (10.0 / 41.0) * tpi2 / pow(wk, 16.0 / 3.0 - kap) * ck / (3.0 * kap - 1.0) * g53 * hyp;
    T highA = 2.0 * 17.493418327624862 / 3.0 * ck / pow(wk, 5.0) * (2.0 * gk2 / (2.0 + kap) - 1.0);   // 2 pi^(5/2)/3
    T NhighA = highA * (pow(3.0 / kap, 4.75) + 0.6) + small;
    T xA = pow(-1.75 + 1.6 * kap, -0.86);
    K.v[KK_LOWA] = lowA;
    K.v[KK_RA] = pow(lowA / NhighA, xA);
    K.v[KK_EA] = xA * (3.0 * kap - 1.0) / 6.0;
    K.v[KK_MA] = -1.0 / xA;
    T NlowQA = -(25.0 / 48.0) * lowA;
    T NhighQA = -(441.0 * pow(kap, -5.76) + 0.55) * highA + small;
    T xQA = 1.4 * pow(kap, -1.15);
    K.v[KK_NLOWQA] = NlowQA;
    K.v[KK_RQA] = pow(NlowQA / NhighQA, xQA);
    K.v[KK_EQA] = xQA * (3.0 * kap - 1.0) / 6.0;
    K.v[KK_MQA] = -1.0 / xQA;
    K.v[KK_CVA1] = -(0.77 / w) * pow(kap, -0.7) * lowA;
    K.v[KK_CVA2] = -(14.3 * pow(w, -0.928)) * (169.0 * pow(kap, -8.0) + 13.0 / 2500.0 * kap - 263.0 / 5000.0 + 47.0 / (200.0 * kap)) * highA;
    T xVA = 1.22 * pow(kap, -1.136) + 0.007;
    K.v[KK_XVA] = xVA;
    K.v[KK_EVA] = xVA * (3.0 * kap - 1.0) / 6.0;
    K.v[KK_MVA] = -1.0 / xVA;
    // rotativities (Marszewski+ 2021), linear in kappa between the fits at kappa = 3.5, 4, 4.5, 5: the fits'
    // w-dependent factors times the interpolation weights
    T kc = dmin(dmax(kap, 3.5), 5.0);
    const int i = val(kc) < 4.0 ? 1 : (val(kc) < 4.5 ? 2 : 3);
    const double klo = 3.5 + (i - 1) * 0.5;
    T f = (kc - klo) * 2.0;
    T sw = sqrt(w), e5 = exp(-5.0 * w);
    auto aq = [&](int n) -> T {
        if (n == 1) return 17.0 * w - 3.0 * sw + 7.0 * sw * e5;
        if (n == 2) return 46.0 * w / 3.0 - 5.0 * sw / 3.0 + 17.0 * sw / 3.0 * e5;
        if (n == 3) return 14.0 * w - 13.0 * sw / 8.0 + 4.5 * sw * e5;
        return 12.5 * w - sw + 5.0 * sw * e5;
    };
    auto rv = [&](int n) -> T {
        if (n == 1) return (w * w + 2.0 * w + 1.0) / (3.125 * w * w + 4.0 * w + 1.0);
        if (n == 2) return (w * w + 54.0 * w + 50.0) / (30.0 * w * w / 11.0 + 134.0 * w + 50.0);
        if (n == 3) return (w * w + 43.0 * w + 38.0) / (7.0 * w * w / 3.0 + 92.5 * w + 38.0);
        return (w + 13.0 / 14.0) / (2.0 * w + 13.0 / 14.0);
    };
    K.v[KK_FQ0] = (1.0 - f) * aq(i);
    K.v[KK_FQ1] = f * aq(i + 1);
    K.v[KK_FV0] = (1.0 - f) * rv(i);
    K.v[KK_FV1] = f * rv(i + 1);
    K.v[KK_RHOQ] = EE * EE * nuc * nuc / (ME * CL);
    K.v[KK_RHOV] = 2.0 * EE * EE * nuc / (ME * CL) * gK[5];
    K.v[KK_IFIT] = T((double)i);
}
// the X-dependent factors of the rotativity fits n = 1..4 (kappa.jl:72-97); p084 = X^0.84, xm = X^-1/2
template <class T>
KS_DEV T kappa_rhoQ_x(int n, T X, T p084) {
    if (n == 1) return 1.0 - exp(-p084 / 30.0) - sin(X / 10.0) * exp(-1.5 * pow(X, 0.471));
    if (n == 2) return 1.0 - exp(-p084 / 18.0) - sin(X / 6.0) * exp(-1.75 * sqrt(X));
    if (n == 3) return 1.0 - exp(-p084 / 12.0) - sin(X / 4.0) * exp(-2.0 * pow(X, 0.525));
    return 1.0 - exp(-p084 / 8.0) - sin(0.375 * X) * exp(-2.25 * pow(X, 0.541));
}
template <class T>
KS_DEV T kappa_rhoV_x(int n, T xm) {
    const double b = n == 1 ? 0.447 : (n == 2 ? 0.391 : (n == 3 ? 0.348 : 0.313));
    return 1.0 - 0.17 * log1p(b * xm);
}
// One sample's kappa coefficients from the splat's factors: K(f) gives factor f (a value, or a dual carrying its
// partials in the splat's rows); ne the density, nu the fluid-frame frequency, (sth, cth) the pitch angle.
template <class T, class KF>
KS_DEV CoefT<T> kappa_coefs(const KF& K, T ne, T nu, T sth, T cth) {
    CoefT<T> c;
    if (!(sth > 0.0)) sth = T(sin_floor(sth));
    const double small = 1e-40;
    T X = nu / (K(KK_NUW1) * sth);
    T pref = ne * K(KK_JPRE) * sth;
    T X13 = cbrt(X);
    T pXI = pow(X, K(KK_EI));
    c.jI = pref * K(KK_NLOWI) * X13 * pow(1.0 + pXI * K(KK_RI), K(KK_MI));
    c.jQ = -(pref * K(KK_NLOWQ) * X13 * pow(1.0 + pow(X, K(KK_EQ)) * K(KK_RQ), K(KK_MQ)));
    // the sin^-p - 1 factors from one logarithm of sin^2 (sin_pow_m1)
    T ls2 = cth * cth < 0.5 ? log1p(-(cth * cth)) : 2.0 * log(sth);
    T Xm035 = pow(X, -0.35), sX = sqrt(X);
    T NlowV = K(KK_CV1) * pow_nonneg(expm1(-1.2 * ls2), 0.48) * Xm035;
    T NhighV = K(KK_CV2) * pow_nonneg(expm1(-1.25 * ls2), 0.44) / sX - small;
    const double sgn = cth > 0.0 ? 1.0 : (cth < 0.0 ? -1.0 : 0.0);
    c.jV = fabs(NhighV) < small * small ? T(0.0)
         : -(pref * NlowV * X13 * pow(1.0 + pXI * pow(NlowV / NhighV, K(KK_XI)), K(KK_MI))) * sgn;
    T apref = ne * K(KK_APRE) / sth;
    T Xm53 = pow(X, -5.0 / 3.0);
    c.aI = apref * K(KK_LOWA) * Xm53 * pow(1.0 + pow(X, K(KK_EA)) * K(KK_RA), K(KK_MA));
    c.aQ = -(apref * K(KK_NLOWQA) * Xm53 * pow(1.0 + pow(X, K(KK_EQA)) * K(KK_RQA), K(KK_MQA)));
    T NlowVA = K(KK_CVA1) * pow_nonneg(expm1(-1.14 * ls2), 0.446) * Xm035;
    T NhighVA = K(KK_CVA2) * sqrt(expm1(-1.025 * ls2)) / sX + small;
    c.aV = -(apref * NlowVA * Xm53 * pow(1.0 + pow(X, K(KK_EVA)) * pow(NlowVA / NhighVA, K(KK_XVA)), K(KK_MVA))) * sgn;
    const int i = (int)val(K(KK_IFIT));
    T p084 = pow(X, 0.84), xm = 1.0 / sX;
    T fQ = K(KK_FQ0) * kappa_rhoQ_x<T>(i, X, p084) + K(KK_FQ1) * kappa_rhoQ_x<T>(i + 1, X, p084);
    T fV = K(KK_FV0) * kappa_rhoV_x<T>(i, xm) + K(KK_FV1) * kappa_rhoV_x<T>(i + 1, xm);
    T inu = 1.0 / nu;
    c.rQ = -(ne * K(KK_RHOQ) * sth * sth) * inu * inu * inu * fQ;
    c.rV = ne * K(KK_RHOV) * cth * inu * inu * fV;
    return c;
}
template <class T> struct KappaGet {
    const KappaK<T>& k;
    KS_DEV T operator()(int f) const { return k.v[f]; }
};
// the whole evaluation from the splat's population values (the spacetime tangent's path)
template <class T>
KS_DEV CoefT<T> kappa(T ne, T kap, T w, T hyp, const T* gK, T B, T nu, T sth, T cth) {
    KappaK<T> K;
    kappa_consts(kap, w, hyp, gK, B, K);
    return kappa_coefs(KappaGet<T>{K}, ne, nu, sth, cth);
}

// ---- cap, invariants, rotation -----------------------------------------------------------------------------------
KS_DEV double pow2_below(double x) {
    if (!(isfinite(x) && x > 0.0)) return 1.0;
    int e; frexp(x, &e);
    return ldexp(1.0, e - 1);
}
KS_DEV f32 pow2_below(f32 x) {
    if (!(isfinite(x.v) && x.v > 0.0f)) return f32(1.0f);
    int e; frexpf(x.v, &e);
    return f32(ldexpf(1.0f, e - 1));
}
// a power of two has no derivative: the duals' partials stay on f = D / Dp
template <int N, class S> KS_DEV DN<N, S> pow2_below(const DN<N, S>& x) { return DN<N, S>(pow2_below(x.v)); }
template <class T> KS_DEV T cap_factor(T I, T Q, T V) {
    const double fmax = 0.99;
    if (!(I > 0.0)) return (val(Q) == 0.0 && val(V) == 0.0) ? T(1.0) : T(0.0);
    const double s = pow2_below(dmax(dmax(fabs(val(I)), fabs(val(Q))), fabs(val(V))));
    const double is = 1.0 / s;
    T Ih = I * is, q = Q * is, v = V * is;
    T pp = q * q + v * v;
    T p = pp > 0.0 ? sqrt(pp) : T(0.0);
    return p > fmax * Ih ? fmax * Ih / p : T(1.0);
}
// cap_factor and its partials in (I, Q, V): f = fmax I / sqrt(Q^2 + V^2) where the cap acts, else a constant
template <class T> KS_DEV void cap_factor_d(T I, T Q, T V, T& f, T df[3]) {
    const double fmax = 0.99;
    df[0] = df[1] = df[2] = T(0.0);
    if (!(I > 0.0)) { f = (val(Q) == 0.0 && val(V) == 0.0) ? T(1.0) : T(0.0); return; }
    const double s = pow2_below(dmax(dmax(fabs(val(I)), fabs(val(Q))), fabs(val(V))));
    const double is = 1.0 / s;
    T Ih = I * is, q = Q * is, v = V * is;
    T pp = q * q + v * v;
    T p = pp > 0.0 ? sqrt(pp) : T(0.0);
    if (!(p > fmax * Ih)) { f = T(1.0); return; }
    f = fmax * Ih / p;
    df[0] = fmax * is / p;
    df[1] = -(f * q / pp) * is;
    df[2] = -(f * v / pp) * is;
}

// ---- the exact 4x4 step -------------------------------------------------------------------------------------------
template <class T> KS_DEV T phi1(T a) {
    if (fabs(val(a)) < 1e-3) return 1.0 - a * (0.5 - a * (1.0 / 6.0 - a * (1.0 / 24.0 - a / 120.0)));
    return -expm1(-a) / a;
}
// M_n(a) = int_0^1 t^n e^{-at} dt for n = 0..NMAX (step.jl:239-268): for |a| < 1 the 24-term power
// series sum_m t_m / (n + m + 1), t_m = (-a)^m / m!, otherwise the upward recurrence
// M_0 = (1 - e^{-a})/a, M_n = (n M_{n-1} - e^{-a})/a. The series is summed forward from one running
// term t_m for all n at once in FP64 (the reference and the oracle use Horner's rule per n): the same
// truncated series, with 33 reciprocals as constants instead of 24 (NMAX + 1) FP64 Horner coefficients,
// which the compiler would otherwise keep in registers across the sample loops. Its terms fall as
// 1/m!, so the forward sum carries a few ulp of rounding. FP32 keeps Horner's rule.
KS_DEV constexpr double moment_coef(int n, int m) {
    double f = 1.0;
    for (int i = 2; i <= m; ++i) f *= i;
    return ((m & 1) ? -1.0 : 1.0) / (f * (n + m + 1));
}
template <class T> struct base_scalar { using type = T; };
template <int N, class S> struct base_scalar<DN<N, S>> { using type = S; };
template <int NMAX, class T> KS_DEV void moments(T a, T M[10]) {
    if (fabs(val(a)) < 1.0) {
        if constexpr (std::is_same<typename base_scalar<T>::type, double>::value) {
            for (int n = 0; n <= NMAX; ++n) M[n] = T(1.0 / (n + 1));
            T tm = T(1.0);
            for (int m = 1; m <= 23; ++m) {
                tm = tm * (-a) * (1.0 / m);
#pragma unroll
                for (int n = 0; n <= NMAX; ++n) M[n] += tm * (1.0 / (n + m + 1));
            }
        } else {
            // FP32: Horner per n (its coefficients are cheap 32-bit constants)
#pragma unroll
            for (int n = 0; n <= NMAX; ++n) {
                T s = T(moment_coef(n, 23));
#pragma unroll
                for (int m = 22; m >= 0; --m) s = s * a + moment_coef(n, m);
                M[n] = s;
            }
        }
    } else {
        T e = exp(-a);
        M[0] = (1.0 - e) / a;
#pragma unroll
        for (int n = 1; n <= NMAX; ++n) M[n] = ((double)n * M[n - 1] - e) / a;
    }
}
template <class T> KS_DEV T sinhc(T b) {
    if (fabs(val(b)) < 0.1) { T bb = b * b; return 1.0 + bb / 6.0 * (1.0 + bb / 20.0 * (1.0 + bb / 42.0 * (1.0 + bb / 72.0))); }
    return sinh(b) / b;
}
template <class T> KS_DEV T sinc_(T b) {
    if (fabs(val(b)) < 0.1) { T bb = b * b; return 1.0 - bb / 6.0 * (1.0 - bb / 20.0 * (1.0 - bb / 42.0 * (1.0 - bb / 72.0))); }
    return sin(b) / b;
}

// Below this |K' D|^2-sized gate the step takes its cubic Taylor form, whose first dropped term is
// under gate / 24 relative: 1e-17 in FP64; 1e-6 in FP32, which also keeps the general form (its 1/Th
// and the s3 - s2 difference) away from the range where FP32 cancels or overflows.
KS_DEV double taylor_gate(double) { return 1e-17; }
KS_DEV double taylor_gate(f32) { return 1e-6; }
template <int N, class S> KS_DEV double taylor_gate(const DN<N, S>&) { return taylor_gate(S()); }

// The step of one sample (step.jl:164-243) as polynomials in the scaled K' (K = alpha_I + K'):
// O = o0 + o1 K' + o2 K'^2 + o3 K'^3 and Ej = Dp (e0 + e1 K' + e2 K'^2 + e3 K'^3), with
// K'6 = (aQ, aU, aV, rQ, rU, rV) times Dp = pow2_below(D); D > 0. T is the scalar S or duals over it.
// step_coefs_scaled takes the coefficients already multiplied by Dp (exact: Dp is a power of two)
// and f = D / Dp in [1, 2), so duals seeded on the scaled values carry O(1) derivatives.
// The coefficients depend on K' only through its invariants: step_coefs_inv takes the scaled alpha_I, the
// squared norms a2 = |(aQ, aU, aV)|^2, r2 = |(rQ, rU, rV)|^2 and ar = (aQ, aU, aV).(rQ, rU, rV).
template <class T, class S>
KS_DEV void step_coefs_inv(T aI, T a2, T r2, T ar, S f, T o[4], T e4[4]) {
    T a = aI * f;
    T e = exp(-a);
    if (val(a2 + r2) * f * f < 1e-200) {
        T M[10]; moments<1>(a, M);
        o[0] = e; o[1] = -f * e; o[2] = o[3] = T(0.0);
        e4[0] = f * phi1(a); e4[1] = -f * f * M[1]; e4[2] = e4[3] = T(0.0);
        return;
    }
    T h = (a2 - r2) * 0.5;
    T root = sqrt(h * h + ar * ar);
    T L1sq, L2sq;
    if (h >= 0.0) { L1sq = h + root; L2sq = L1sq > 0.0 ? ar * ar / L1sq : T(0.0); }
    else { L2sq = root - h; L1sq = L2sq > 0.0 ? ar * ar / L2sq : T(0.0); }
    T L1 = L1sq > 0.0 ? sqrt(L1sq) : T(0.0), L2 = L2sq > 0.0 ? sqrt(L2sq) : T(0.0);
    T b1 = L1 * f, b2 = L2 * f;
    T Th = L1sq + L2sq;
    if (val(Th) * f * f * (1.0 + val(a2 + r2) * f * f) < taylor_gate(S())) {
        T M[10]; moments<3>(a, M);
        o[0] = e; o[1] = -f * e; o[2] = f * f * 0.5 * e; o[3] = -f * f * f / 6.0 * e;
        e4[0] = f * M[0]; e4[1] = -f * f * M[1]; e4[2] = f * f * f * M[2] * 0.5; e4[3] = -f * f * f * f * M[3] / 6.0;
        return;
    }
    T ech, esh, echm1;
    if (b1 < 1.0) {
        T sh = sinh(0.5 * b1);
        ech = e * cosh(b1); esh = e * sinhc(b1); echm1 = e * 2.0 * sh * sh;
    } else {
        T em = exp(-(a - b1)), ep = exp(-(a + b1));
        ech = 0.5 * (em + ep); esh = (em - ep) / (2.0 * b1); echm1 = 0.5 * (em - 2.0 * e + ep);
    }
    T sb2h = sin(0.5 * b2);
    T versin2 = 2.0 * sb2h * sb2h;
    T cb2 = cos(b2);
    T iTh = 1.0 / Th;
    o[0] = (L2sq * ech + L1sq * e * cb2) * iTh;
    o[2] = (echm1 + e * versin2) * iTh;
    T s3 = f * esh, s2 = f * e * sinc_(b2);
    o[1] = -(s3 * L2sq + s2 * L1sq) * iTh;
    o[3] = -(s3 - s2) * iTh;
    T pa = phi1(a);
    // the moments: M_1 always, M_0..M_9 for the small-b series
    T M[10];
    if (b1 < 0.1 || b2 < 0.1) moments<9>(a, M);
    else moments<1>(a, M);
    // int_0^1 e^{-at}(cosh bt - 1), (1 - cos bt), sin(bt)/b, (sinh(bt)/b - t)
    T icm1, ivs, isc, ish;
    if (b1 < 0.1) {
        T bb = b1 * b1;
        icm1 = bb * (M[2] / 2.0 + bb * (M[4] / 24.0 + bb * (M[6] / 720.0 + bb * M[8] / 40320.0)));
        ish = bb * (M[3] / 6.0 + bb * (M[5] / 120.0 + bb * (M[7] / 5040.0 + bb * M[9] / 362880.0)));
    } else {
        T pm = phi1(a - b1), pp = phi1(a + b1);
        icm1 = 0.5 * (pm + pp) - pa;
        ish = (pm - pp) / (2.0 * b1) - M[1];
    }
    if (b2 < 0.1) {
        T bb = b2 * b2;
        ivs = bb * (M[2] / 2.0 - bb * (M[4] / 24.0 - bb * (M[6] / 720.0 - bb * M[8] / 40320.0)));
        isc = M[1] - bb * (M[3] / 6.0 - bb * (M[5] / 120.0 - bb * (M[7] / 5040.0 - bb * M[9] / 362880.0)));
    } else {
        T sb = sin(b2);
        T den = 1.0 / (a * a + b2 * b2);
        ivs = pa - (a - e * (a * cb2 - b2 * sb)) * den;
        isc = (1.0 - e * (a * sb / b2 + cb2)) * den;
    }
    T Ich = f * (pa + icm1);
    T Ic = f * (pa - ivs);
    T Ish = f * f * (M[1] + ish);
    T Is = f * f * isc;
    e4[0] = (L2sq * Ich + L1sq * Ic) * iTh;
    e4[2] = f * (icm1 + ivs) * iTh;
    e4[1] = -(Ish * L2sq + Is * L1sq) * iTh;
    e4[3] = -(Ish - Is) * iTh;
}
template <class T, class S>
KS_DEV void step_coefs_scaled(const T alS[4], const T rhS[3], S f, T K6[6], T o[4], T e4[4]) {
    const T &aQ = alS[1], &aU = alS[2], &aV = alS[3];
    const T &rQ = rhS[0], &rU = rhS[1], &rV = rhS[2];
    K6[0] = aQ; K6[1] = aU; K6[2] = aV; K6[3] = rQ; K6[4] = rU; K6[5] = rV;
    const T a2 = aQ * aQ + aU * aU + aV * aV, r2 = rQ * rQ + rU * rU + rV * rV;
    const T ar = aQ * rQ + aU * rU + aV * rV;
    step_coefs_inv(alS[0], a2, r2, ar, f, o, e4);
}
template <class T, class S>
KS_DEV void step_coefs(const T al[4], const T rh[3], S D, S& Dp, T K6[6], T o[4], T e4[4]) {
    Dp = pow2_below(D);
    T alS[4], rhS[3];
    for (int i = 0; i < 4; ++i) alS[i] = al[i] * Dp;
    for (int i = 0; i < 3; ++i) rhS[i] = rh[i] * Dp;
    step_coefs_scaled(alS, rhS, S(D / Dp), K6, o, e4);
}

// y = K' v and y = K'^T v with K' = [0 aQ aU aV; aQ 0 rV -rU; aU -rV 0 rQ; aV rU -rQ 0]
template <class T, class U, class R>
KS_DEV void kprime_mul(const T K6[6], const U v[4], R y[4]) {
    const T &aQ = K6[0], &aU = K6[1], &aV = K6[2], &rQ = K6[3], &rU = K6[4], &rV = K6[5];
    y[0] = aQ * v[1] + aU * v[2] + aV * v[3];
    y[1] = aQ * v[0] + rV * v[2] - rU * v[3];
    y[2] = aU * v[0] - rV * v[1] + rQ * v[3];
    y[3] = aV * v[0] + rU * v[1] - rQ * v[2];
}
template <class T, class U, class R>
KS_DEV void kprime_tmul(const T K6[6], const U v[4], R y[4]) {
    const T &aQ = K6[0], &aU = K6[1], &aV = K6[2], &rQ = K6[3], &rU = K6[4], &rV = K6[5];
    y[0] = aQ * v[1] + aU * v[2] + aV * v[3];
    y[1] = aQ * v[0] - rV * v[2] + rU * v[3];
    y[2] = aU * v[0] + rV * v[1] - rQ * v[3];
    y[3] = aV * v[0] - rU * v[1] + rQ * v[2];
}

// O and E = Ej j for K = alpha_I + K' (step.jl:164-243), D > 0. Writes O row-major.
template <class S>
KS_DEV void transfer_step(const S j4[4], const S al[4], const S rh[3], S D, S O[16], S E[4]) {
    S Dp, K6[6], o[4], e4[4];
    step_coefs(al, rh, D, Dp, K6, o, e4);
    const S aQ = K6[0], aU = K6[1], aV = K6[2], rQ = K6[3], rU = K6[4], rV = K6[5];
    const S zero(0.0);
    const S K[16] = {zero, aQ, aU, aV, aQ, zero, rV, -rU, aU, -rV, zero, rQ, aV, rU, -rQ, zero};
    S K2[16], K3[16];
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) {
        S s(0.0);
        for (int q = 0; q < 4; ++q) s += K[4 * i + q] * K[4 * q + k];
        K2[4 * i + k] = s;
    }
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) {
        S s(0.0);
        for (int q = 0; q < 4; ++q) s += K2[4 * i + q] * K[4 * q + k];
        K3[4 * i + k] = s;
    }
    for (int i = 0; i < 4; ++i) {
        S ei(0.0);
        for (int k = 0; k < 4; ++k) {
            const double id = i == k ? 1.0 : 0.0;
            int ix = 4 * i + k;
            O[ix] = o[0] * id + o[1] * K[ix] + o[2] * K2[ix] + o[3] * K3[ix];
            S ej = e4[0] * id + e4[1] * K[ix] + e4[2] * K2[ix] + e4[3] * K3[ix];
            ei += ej * j4[k];
        }
        E[i] = Dp * ei;
    }
}

// T_out = E + O T_in for one sample from its polynomial coefficients (the tails recursion,
// dual_gradient.jl:129-147), by repeated products with K' instead of the 4x4 matrices
template <class S>
KS_DEV void step_apply(const S j4[4], const S al[4], const S rh[3], S D, const S Tin[4], S Tout[4]) {
    S Dp, K6[6], o[4], e4[4];
    step_coefs(al, rh, D, Dp, K6, o, e4);
    S v[4], u[4], y[4], z[4];
    for (int i = 0; i < 4; ++i) { v[i] = Tin[i]; u[i] = j4[i]; Tout[i] = o[0] * Tin[i] + Dp * e4[0] * j4[i]; }
    for (int n = 1; n < 4; ++n) {
        kprime_mul(K6, v, y); kprime_mul(K6, u, z);
        for (int i = 0; i < 4; ++i) { v[i] = y[i]; u[i] = z[i]; Tout[i] += o[n] * v[i] + Dp * e4[n] * u[i]; }
    }
}

// One FP32 sample's step inputs passed by value to the out-of-line FP64 steps of the FP32 step adjoints: an array
// whose address reaches an out-of-line call lives in scratch memory for the whole caller, so the caller's records
// are copied into this only on the rare path that makes the call (the FP32 step adjoints: 20% faster).
struct Step32 { f32 j[4], a[4], r[3], D; };
KS_DEV Step32 step32_of(const f32 j4[4], const f32 a4[4], const f32 r3[3], f32 D) {
    Step32 s;
    for (int i = 0; i < 4; ++i) { s.j[i] = j4[i]; s.a[i] = a4[i]; }
    for (int i = 0; i < 3; ++i) s.r[i] = r3[i];
    s.D = D;
    return s;
}

}  // namespace ksd
