// Gaussian splat models and the fused-march consumers for the CPU oracle (test infrastructure; see
// ks_scalar.hpp). Restates src/Splats/Splats.jl, polarized.jl, populations.jl, trajectories.jl and
// the RadiativeTransport consumer of src/Transfer/transport.jl.
//
// Parameter matrices are column-major `nrows x ns` exactly as in Julia; row indices below are
// 0-based (Julia row r is index r - 1 here).
//   thin (14 rows):       x y z s1 s2 s3 q1..q4 t0 logw logA omega          (Splats.jl:54)
//   thermal (21 rows):    x y z s1 s2 s3 q1..q4 t0 logw logne logTe logB thB phB u1 u2 u3 omega
//                                                                          (polarized.jl:17-19)
//   power law (22 rows):  ... logne p loggmin logB thB phB u1 u2 u3 omega (populations.jl:13-14)
//   kappa (22 rows):      ... logne kappa logw_kappa logB thB phB u1 u2 u3 omega (populations.jl:24-25)
//   knots: thermal rows, centre interpolated from 3 x nknots x ns knots    (trajectories.jl:80-128)
#pragma once
#include "ks_geodesic.hpp"
#include "ks_transfer.hpp"
#include "ks_dual.hpp"

namespace kso {

enum ModelKind : int { MODEL_THIN = 0, MODEL_THERMAL = 1, MODEL_POWERLAW = 2, MODEL_KAPPA = 3, MODEL_KNOT = 4 };

struct Model {
    int kind = MODEL_THERMAL;
    int nrows = 21, ns = 0;
    const double* params = nullptr;   // nrows x ns, column-major
    const double* hyp = nullptr;      // kappa: 2F1 factor per splat
    double gamma_max = 1e5;           // power law
    const double* knots = nullptr;    // 3 x nknots x ns
    const double* knot_times = nullptr;
    int nknots = 0;
    // per-splat Bessel values that depend only on the splat (K0, K1, K2 at 1/Theta_e for thermal
    // and knot splats, at 1/w for kappa splats), filled once per model
    const double* bessel = nullptr;
    // CompositeModel (transport.jl:298-316): the elements of this model are followed by those of
    // `next` (and its successors); coefficients of overlapping elements add
    const Model* next = nullptr;
    double P(int row, int i) const { return params[row + nrows * i]; }
    // centres that follow knots (KnotSplats, trajectories.jl:80-128; here for any fluid population)
    bool has_knots() const { return knots != nullptr && knot_times != nullptr && nknots > 0 && kind != MODEL_THIN; }
};

// pattern-frame offset of (x, y, z) from the centre of splat i at time t (Splats.jl:58-69)
template <class T> void quat_rot(T q1, T q2, T q3, T q4, T R[3][3]) {
    T n = ksqrt(q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4);
    T w = q1 / n, x = q2 / n, y = q3 / n, z = q4 / n;
    R[0][0] = 1.0 - 2.0 * (y * y + z * z); R[0][1] = 2.0 * (x * y - w * z); R[0][2] = 2.0 * (x * z + w * y);
    R[1][0] = 2.0 * (x * y + w * z); R[1][1] = 1.0 - 2.0 * (x * x + z * z); R[1][2] = 2.0 * (y * z - w * x);
    R[2][0] = 2.0 * (x * z - w * y); R[2][1] = 2.0 * (y * z + w * x); R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}
template <class T> void pattern_offset(const Model& m, int i, T t, T x, T y, T z, T u[3]) {
    int iw = m.nrows - 1;
    T ph = m.P(iw, i) * (t - m.P(10, i));
    T sp = ksin(ph), cp = kcos(ph);
    T cx = m.P(0, i) * cp - m.P(1, i) * sp;
    T cy = m.P(0, i) * sp + m.P(1, i) * cp;
    T d[3] = {x - cx, y - cy, z - m.P(2, i)};
    T dr[3] = {d[0] * cp + d[1] * sp, -d[0] * sp + d[1] * cp, d[2]};
    T R[3][3]; quat_rot(T(m.P(6, i)), T(m.P(7, i)), T(m.P(8, i)), T(m.P(9, i)), R);
    for (int a = 0; a < 3; ++a) u[a] = R[0][a] * dr[0] + R[1][a] * dr[1] + R[2][a] * dr[2];
}
template <class T> T gauss_exponent(const Model& m, int i, T t, const T u[3]) {
    T q2 = ksq(u[0] * kexp(T(-m.P(3, i)))) + ksq(u[1] * kexp(T(-m.P(4, i)))) + ksq(u[2] * kexp(T(-m.P(5, i))));
    T tau2 = ksq((t - m.P(10, i)) * kexp(T(-m.P(11, i))));
    return q2 + tau2;
}
template <class T> T splat_weight(const Model& m, int i, T t, T x, T y, T z) {
    T u[3]; pattern_offset(m, i, t, x, y, z, u);
    return kexp(-gauss_exponent(m, i, t, u) / 2.0);
}
// bounding-sphere cull (polarized.jl:110-120)
template <class T> bool outside_support(const Model& m, int i, T t, T x, T y, T z) {
    int iw = m.nrows - 1;
    T ph = m.P(iw, i) * (t - m.P(10, i));
    T sp = ksin(ph), cp = kcos(ph);
    T cx = m.P(0, i) * cp - m.P(1, i) * sp;
    T cy = m.P(0, i) * sp + m.P(1, i) * cp;
    T d2 = ksq(x - cx) + ksq(y - cy) + ksq(z - m.P(2, i));
    double smax = std::max(std::max(m.P(3, i), m.P(4, i)), m.P(5, i));
    return d2 > support_radius2() * std::exp(2.0 * smax);
}
// knot centre (trajectories.jl:93-111)
template <class T> void knot_centre(const Model& m, int i, T t, T c[3]) {
    const double* ts = m.knot_times;
    int n = m.nknots;
    const double* K = m.knots + 3 * n * i;
    if (t <= ts[0]) { for (int a = 0; a < 3; ++a) c[a] = K[a]; return; }
    if (t >= ts[n - 1]) { for (int a = 0; a < 3; ++a) c[a] = K[3 * (n - 1) + a]; return; }
    int k = 0;
    while (k < n - 2 && ts[k + 1] <= t) ++k;
    T w = (t - ts[k]) / (ts[k + 1] - ts[k]);
    for (int a = 0; a < 3; ++a) c[a] = (1.0 - w) * K[3 * k + a] + w * K[3 * (k + 1) + a];
}

// ---- one element's fluid coefficients and frame -----------------------------------------------------
// the ray's spacetime and screen constants; R = double, or a dual scalar for the spacetime tangent
template <class R> struct RayCtxT { R a, alpha, beta, thetao, eta, lam; };
using RayCtx = RayCtxT<double>;

// K0, K1, K2 at x_b = 1 / Theta_e (or 1 / w) from the splat's table kb at the primal value: the table itself
// for plain scalars, and for dual numbers the table with the derivatives K0' = -K1, K1' = -K0 - K1/x,
// K2' = K0' + 2 K1'/x - 2 K1/x^2 carried through x_b
template <class T> struct is_dual : std::false_type {};
template <int N, class S> struct is_dual<Dual<N, S>> : std::true_type {};
template <class T> void bessel_at(const double* kb, const T& xb, T& k0, T& k1, T& k2) {
    if constexpr (is_dual<T>::value) {
        const double x = val(xb);
        const double d0 = -kb[1], d1 = -kb[0] - kb[1] / x, d2 = d0 + 2.0 * d1 / x - 2.0 * kb[1] / (x * x);
        k0 = chain(kb[0], d0, xb); k1 = chain(kb[1], d1, xb); k2 = chain(kb[2], d2, xb);
    } else {
        k0 = T(kb[0]); k1 = T(kb[1]); k2 = T(kb[2]);
    }
}

// The Gaussian exponent of splat i at (t, x, y, z) from its rows pr (given in the scalar T, so that dual
// rows carry their derivatives): the pattern-frame offset rotated by omega (t - t0) about the spin axis
// (Splats.jl:58-93), or for a model with knots the offset from the knot-interpolated centre (trajectories.jl:
// 93-128; no pattern rotation), where the centre rows translate the trajectory rigidly (their derivative is
// the derivative with respect to a shift of every knot).
template <class T> T gexp_rows(const Model& m, int i, const T* pr, T t, T x, T y, T z) {
    T dr[3];
    if (m.has_knots()) {
        T c[3]; knot_centre(m, i, t, c);
        for (int a = 0; a < 3; ++a) c[a] = c[a] + (pr[a] - m.P(a, i));
        dr[0] = x - c[0]; dr[1] = y - c[1]; dr[2] = z - c[2];
    } else {
        T ph = pr[m.nrows - 1] * (t - pr[10]);
        T sp = ksin(ph), cp = kcos(ph);
        T cx = pr[0] * cp - pr[1] * sp;
        T cy = pr[0] * sp + pr[1] * cp;
        T d[3] = {x - cx, y - cy, z - pr[2]};
        dr[0] = d[0] * cp + d[1] * sp; dr[1] = -d[0] * sp + d[1] * cp; dr[2] = d[2];
    }
    T R[3][3]; quat_rot(pr[6], pr[7], pr[8], pr[9], R);
    T u[3];
    for (int a = 0; a < 3; ++a) u[a] = R[0][a] * dr[0] + R[1][a] * dr[1] + R[2][a] * dr[2];
    T q2 = ksq(u[0] * kexp(-pr[3])) + ksq(u[1] * kexp(-pr[4])) + ksq(u[2] * kexp(-pr[5]));
    T tau2 = ksq((t - pr[10]) * kexp(-pr[11]));
    return q2 + tau2;
}

// One element's fluid coefficients and frame at a sample (element(::PolarizedSplats / PowerLawSplats /
// KappaSplats / KnotSplats), polarized.jl:96-150, populations.jl:13-104, trajectories.jl:113-128), from the
// splat's rows pr in the scalar T. Plain scalars take the per-splat tables (Bessel values, the kappa 2F1
// factor) as they are; dual rows recompute what depends on them (Gamma functions, 2F1) or carry the tables'
// derivatives, so that the result is differentiable in every row.
template <class T, class RS> bool element_rows(const Model& m, int i, const RayCtxT<RS>& rc, const Sample<T>& s, T t_obs, T nu_obs,
                                               const T* pr, Coef<T>& cf, Frame<T>& fr) {
    T x, y, z;
    quasi_cartesian_ks(T(rc.a), s.r, s.th, s.ph, x, y, z);
    T t = t_obs - s.t;
    if (!m.has_knots()) {
        if constexpr (is_dual<T>::value) { if (outside_support(m, i, val(t), val(x), val(y), val(z))) return false; }
        else if (outside_support(m, i, t, x, y, z)) return false;
    }
    T G = kexp(-gexp_rows(m, i, pr, t, x, y, z) / 2.0);
    if (!(G > WEIGHT_CUTOFF)) return false;
    const bool th_rows = m.kind == MODEL_THERMAL || m.kind == MODEL_KNOT;
    const int rB = th_rows ? 14 : 15, rthB = rB + 1, rphB = rB + 2, ru = rB + 3;
    T Bmag = kexp(pr[rB]);
    T st = ksin(pr[rthB]), ct = kcos(pr[rthB]), sp = ksin(pr[rphB]), cp = kcos(pr[rphB]);
    T B[3] = {Bmag * st * cp, Bmag * st * sp, Bmag * ct};
    T u[3] = {pr[ru], pr[ru + 1], pr[ru + 2]};
    fr = local_frame(T(rc.a), s.r, s.th, T(rc.eta), T(rc.lam), s.nur, s.nuth, T(rc.alpha), T(rc.beta), T(rc.thetao), u, B);
    T nuf = nu_obs / fr.g;
    T thB = kacos(kmin(kmax(fr.cosB, T(-1.0)), T(1.0)));
    T ne = kexp(pr[12]) * G;
    const double* kb = m.bessel + 3 * i;
    if (th_rows) {
        T Th = kexp(pr[13]);
        T k0, k1, k2;
        bessel_at(kb, T(1.0) / Th, k0, k1, k2);
        cf = thermal_synchrotron_k(ne, Th, Bmag, nuf, thB, k0, k1, k2);
    } else if (m.kind == MODEL_POWERLAW) {
        cf = powerlaw_synchrotron(ne, pr[13], kexp(pr[14]), T(m.gamma_max), Bmag, nuf, thB);
    } else {
        T w = kexp(pr[14]);
        T hyp, k0k2;
        if constexpr (is_dual<T>::value) {
            hyp = kappa_hypergeometric_t(pr[13], w);
            T k0, k1, k2;
            bessel_at(kb, T(1.0) / w, k0, k1, k2);
            k0k2 = k0 / (k2 + 1e-40);
        } else {
            hyp = T(m.hyp[i]);
            k0k2 = T(kb[0] / (kb[2] + 1e-40));
        }
        cf = kappa_synchrotron(ne, pr[13], w, Bmag, nuf, thB, hyp, k0k2);
    }
    return true;
}

template <class T, class RS> bool element(const Model& m, int i, const RayCtxT<RS>& rc, const Sample<T>& s, T t_obs, T nu_obs, Coef<T>& cf, Frame<T>& fr) {
    T pr[32];
    for (int r = 0; r < m.nrows; ++r) pr[r] = T(m.P(r, i));
    return element_rows(m, i, rc, s, t_obs, nu_obs, pr, cf, fr);
}

// screen-basis invariant coefficients of one element (transport.jl:223-228)
template <class T> bool screen_coefficients(const Coef<T>& cf, const Frame<T>& fr, T nu_obs, Screen<T>& sc) {
    if (!(cf.jI > 0.0 || cf.aI > 0.0 || val(cf.rQ) != 0.0 || val(cf.rV) != 0.0)) return false;
    Coef<T> ci = invariants(cap_polarization(cf), nu_obs / fr.g);
    sc = rotate_to_screen(ci, fr.chi);
    return true;
}

// ---- radiative state (step.jl:342-360) --------------------------------------------------------------------
template <class T> struct RadState { M4<T> P; T S[4]; };
template <class T> RadState<T> radstate_zero() { RadState<T> s; s.P = m4_eye<T>(); for (int i = 0; i < 4; ++i) s.S[i] = 0.0; return s; }

// the winding counter (step.jl:362-398)
template <class T> struct Winding { T zprev; int n; bool inside, crossed; };
template <class T> Winding<T> winding_zero() { return {T(NAN), 0, false, false}; }
template <class T> void wind(Winding<T>& w, T z, T h) {
    if (kisnan(w.zprev)) { w.zprev = z; w.inside = kabs(z) < h; w.crossed = false; return; }
    bool inside_now = kabs(z) < h;
    bool crossing = w.zprev * z < 0.0;
    if (w.inside || inside_now) {
        w.crossed = w.crossed || crossing;
        if (w.inside && !inside_now) { if (w.crossed) w.n += 1; w.crossed = false; }
        else if (!w.inside && inside_now) { w.crossed = crossing; }
        w.inside = inside_now;
    } else if (crossing) {
        w.n += 1;
    }
    w.zprev = z;
}

// sum of all elements' screen coefficients at a sample (transport.jl:254-276); returns whether any
// element is active. A composite model visits its parts in order, each part's elements in order
// (element(::CompositeModel), transport.jl:312-316), so the sum runs over the same sequence.
template <class T, class RS> bool accumulate_elements(const Model& m0, const RayCtxT<RS>& rc, const Sample<T>& s, T t_obs, T nu_obs, T j4[4], T a4[4], T r3[3]) {
    for (int q = 0; q < 4; ++q) { j4[q] = 0.0; a4[q] = 0.0; }
    for (int q = 0; q < 3; ++q) r3[q] = 0.0;
    bool active = false;
    for (const Model* pm = &m0; pm; pm = pm->next) {
        const Model& m = *pm;
        for (int i = 0; i < m.ns; ++i) {
            Coef<T> cf; Frame<T> fr;
            if (!element(m, i, rc, s, t_obs, nu_obs, cf, fr)) continue;
            Screen<T> sc;
            if (!screen_coefficients(cf, fr, nu_obs, sc)) continue;
            for (int q = 0; q < 4; ++q) { j4[q] = j4[q] + sc.j[q]; a4[q] = a4[q] + sc.a[q]; }
            for (int q = 0; q < 3; ++q) r3[q] = r3[q] + sc.r[q];
            active = true;
        }
    }
    return active;
}

// one sample of the polarized transport (transport.jl:263-290); returns true when a step was taken
template <class T, class RS> bool sample_step(const Model& m, const RayCtxT<RS>& rc, const Sample<T>& s, T dtau, T t_obs, T nu_obs, T L, M4<T>& O, T E[4]) {
    T a = T(rc.a);
    const double av = val(rc.a);
    T rh = 1.0 + std::sqrt(1.0 - av * av);
    if (!(s.ok && s.r > rh * (1.0 + 1e-3))) return false;
    T j4[4], a4[4], r3[3];
    if (!accumulate_elements(m, rc, s, t_obs, nu_obs, j4, a4, r3)) return false;
    T c = kcos(s.th);
    T Sg = s.r * s.r + a * a * c * c;
    T D = L / (nu_obs / NU0) * Sg * dtau;
    M4<T> Ej;
    step_operator(a4, r3, D, O, Ej);
    for (int i = 0; i < 4; ++i) E[i] = Ej.m[i][0] * j4[0] + Ej.m[i][1] * j4[1] + Ej.m[i][2] * j4[2] + Ej.m[i][3] * j4[3];
    return true;
}
template <class T> void advance(RadState<T>& st, const M4<T>& O, const T E[4]) {
    T PE[4];
    for (int i = 0; i < 4; ++i) PE[i] = st.P.m[i][0] * E[0] + st.P.m[i][1] * E[1] + st.P.m[i][2] * E[2] + st.P.m[i][3] * E[3];
    for (int i = 0; i < 4; ++i) st.S[i] = st.S[i] + PE[i];
    st.P = m4_mul(st.P, O);
}

// ---- thin renderer (Splats.jl:113-128) ---------------------------------------------------------------------------
template <class T> T thin_sample(const Model& m, const RayCtx& rc, const Sample<T>& s, T dtau, T t_obs) {
    T a = rc.a;
    T rh = 1.0 + std::sqrt(1.0 - rc.a * rc.a);
    if (!(s.ok && rh * (1.0 + 1e-3) < s.r && s.r < 1e3)) return 0.0;
    T x, y, z;
    quasi_cartesian_ks(T(a), s.r, s.th, s.ph, x, y, z);
    T t = t_obs - s.t;
    T jt = 0.0;
    for (int i = 0; i < m.ns; ++i) {
        T u[3]; pattern_offset(m, i, t, x, y, z, u);
        jt = jt + kexp(T(m.P(12, i)) - gauss_exponent(m, i, t, u) / 2.0);
    }
    T g = zamo_redshift(T(a), s.r, s.th, T(rc.lam));
    T c = kcos(s.th);
    T Sg = s.r * s.r + a * a * c * c;
    return g * g * jt * Sg * dtau;
}

}  // namespace kso
