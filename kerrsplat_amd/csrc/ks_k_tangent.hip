// The spacetime tangent (spacetime_valgrad / spacetime_jacobian, joint.jl:74-154): the observed Stokes vector
// of every (pixel, time, frequency) item and its forward-mode partials with respect to x = (a, theta_o, ln L),
// carried as three-partial duals (DN<3, double>) through the pixel constants (K1), the closed-form ("Direct")
// march and the polarized transport. The geometry code is the FP64 code of the image kernels instantiated on
// the dual scalar; one thread marches and transports one item (the reference's Dual{3} cache costs about
// four forward passes and is used by the spacetime block of a fit, not on the image path).
#include "ks_kernels.hpp"

namespace ksd {

using D3 = DN<3, double>;

// a splat's constants as duals with zero partials (they do not depend on the spacetime)
KS_DEV SplatCT<D3> splat_dual(const SplatC& s) {
    SplatCT<D3> o;
    o.x = D3(s.x); o.y = D3(s.y); o.z = D3(s.z);
    for (int i = 0; i < 9; ++i) o.R[i] = D3(s.R[i]);
    for (int i = 0; i < 3; ++i) { o.is[i] = D3(s.is[i]); o.bh[i] = D3(s.bh[i]); o.u[i] = D3(s.u[i]); }
    o.t0 = D3(s.t0); o.iw = D3(s.iw); o.omega = D3(s.omega); o.rsup2 = D3(s.rsup2); o.lnA = D3(s.lnA); o.Bmag = D3(s.Bmag);
    o.p1 = D3(s.p1); o.p2 = D3(s.p2); o.p3 = D3(s.p3);
    for (int i = 0; i < 6; ++i) o.g[i] = D3(s.g[i]);
    o.knot0 = s.knot0; o.nk = s.nk; o.kt0 = s.kt0; o.kind = s.kind;
    return o;
}

// The splat sum of one sample in duals (accumulate_elements, transport.jl:254-276; the same steps as the
// image kernels' sample_coefficients, with the geometry built in the dual scalar).
template <int KIND>
KS_DEV bool tangent_coefficients(const SplatC* __restrict__ sp, int ns, const D3& a, const D3& alpha, const D3& beta, const D3& tsin,
                                 const D3& eta, const D3& lam, double t_obs, double nu, const SampT<D3>& s, const D3& x, const D3& y,
                                 const D3& z, D3 j4[4], D3 a4[4], D3 r3[3]) {
    const D3 t = t_obs - s.t;
    for (int i = 0; i < 4; ++i) { j4[i] = D3(0.0); a4[i] = D3(0.0); }
    r3[0] = r3[1] = r3[2] = D3(0.0);
    bool active = false, geo_ready = false;
    GeoT<D3> G;
    for (int i = 0; i < ns; ++i) {
        const SplatCT<D3> sc = splat_dual(sp[i]);
        D3 sph, cph;
        if (outside_support(sc, t, x, y, z, sph, cph)) continue;
        const D3 gexp = gauss_exponent(sc, t, x, y, z, sph, cph);
        const D3 G_ = exp(-0.5 * gexp);
        if (!(G_ > WEIGHT_CUTOFF)) continue;
        if (!geo_ready) { make_geo(a, s.r, s.th, eta, lam, s.nur, s.nuth, G); geo_ready = true; }
        const FrameT<D3> fr = local_frame(G, a, alpha, beta, tsin, sc.u, sc.bh);
        const D3 nuf = nu / fr.g;
        const D3 cb = dmin(dmax(fr.cosB, -1.0), 1.0);
        D3 sth, cth;
        pitch_sincos(cb, sth, cth);
        const D3 ne = exp(sc.lnA) * G_;
        CoefT<D3> cf;
        const int kind = KIND == MODEL_COMPOSITE ? sc.kind : KIND;
        if (kind == MODEL_THERMAL) cf = thermal(ne, sc.p1, sc.p2, sc.p3, sc.Bmag, nuf, sth, cth);
        else if (kind == MODEL_POWERLAW) cf = powerlaw(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        else cf = kappa(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        if (!(cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ != 0.0 || cf.rV != 0.0)) continue;
        const D3 fe = cap_factor(cf.jI, cf.jQ, cf.jV), fa = cap_factor(cf.aI, cf.aQ, cf.aV);
        const D3 nh = nuf * (1.0 / NU0);
        const D3 i2 = 1.0 / (nh * nh);
        D3 s2c, c2c; sincos(2.0 * fr.chi, &s2c, &c2c);
        const D3 jQ = fe * cf.jQ * i2, aQ = fa * cf.aQ * nh, rQ = cf.rQ * nh;
        j4[0] += cf.jI * i2; j4[1] += jQ * c2c; j4[2] += jQ * s2c; j4[3] += fe * cf.jV * i2;
        a4[0] += cf.aI * nh; a4[1] += aQ * c2c; a4[2] += aQ * s2c; a4[3] += fa * cf.aV * nh;
        r3[0] += rQ * c2c; r3[1] += rQ * s2c; r3[2] += cf.rV * nh;
        active = true;
    }
    return active;
}

template <int KIND>
KS_DEV void tangent_item(int64_t item, double a, double thetao, double lnL, const double* __restrict__ al, const double* __restrict__ be,
                         int64_t npix, int N, const SplatC* __restrict__ sp, int ns, const double* __restrict__ times, int nt,
                         const double* __restrict__ nus, int nmax, double slab, const double* glx, const double* glw,
                         double* __restrict__ out, double* __restrict__ dout) {
    const int64_t j = item % npix;
    const int it = (int)((item / npix) % nt), inu = (int)(item / (npix * nt));
    const double nu = nus[inu], t_obs = times[it];
    const D3 ad = D3::seed(a, 0), td = D3::seed(thetao, 1), Ld = exp(D3::seed(lnL, 2));
    const SpinT<D3> S = make_spin(ad);
    const PixCT<D3> p = pixel_constants(S, td, D3(al[j]), D3(be[j]), glx, glw);
    const VisCT<D3> vis = vis_constants(p, ad, td);
    const D3 tsin = sin(td);
    const D3 dtau = p.tau_total / (double)(N + 1);
    const double rp = val(S.rp);
    WindF wf{rp, slab, nmax, N + 1, NAN, 0, false, false};
    SegMapT<D3> sm;
    sm.init();
    for (int k = 1; k <= N; ++k) {
        const SampT<D3> s = direct_sample(p, S, vis, dtau * (double)k);
        if (nmax >= 0) {
            const Samp sv{0.0, val(s.r), val(s.th), 0.0, s.nur, s.nuth, s.ok};
            wf(k, sv, 0.0);
            if (wf.kstop <= k) break;
        }
        if (!(s.ok && s.r > rp * (1.0 + 1e-3))) continue;
        D3 x, y, z;
        quasi_cartesian_ks(ad, S.rp, S.rm, s.r, s.th, s.ph, x, y, z);
        D3 j4[4], a4[4], r3[3];
        if (!tangent_coefficients<KIND>(sp, ns, ad, p.alpha, p.beta, tsin, p.eta, p.lam, t_obs, nu, s, x, y, z, j4, a4, r3)) continue;
        const D3 ct = cos(s.th);
        const D3 D = Ld / (nu * (1.0 / NU0)) * (s.r * s.r + ad * ad * ct * ct) * dtau;
        sm.step(j4, a4, r3, D);
    }
    const double nh = nu * (1.0 / NU0), nh3 = nh * nh * nh;
    for (int q = 0; q < 4; ++q) {
        out[4 * item + q] = sm.Sv[q].v * nh3;
        for (int k = 0; k < 3; ++k) dout[12 * item + 3 * q + k] = sm.Sv[q].d[k] * nh3;
    }
}

// One thread per item (j + npix (it + nt inu)), pixels in the cache's screen order; out: 4 per item,
// dout: 12 per item (Stokes q, partial k at 3 q + k). The samples are the Direct marcher's: tau_k = k dtau,
// dtau = tau_total / (N + 1), k = 1..N (direct_march.jl:27-124); nmax >= 0 cuts each ray after its
// nmax-th half-orbit passage through the slab (the counter runs on the values).
template <int KIND>
__global__ void __launch_bounds__(64) ks_tangent(double a, double thetao, double lnL, const double* __restrict__ al,
                                                 const double* __restrict__ be, int64_t npix, int N, const SplatC* __restrict__ sp, int ns,
                                                 const double* __restrict__ times, int nt, const double* __restrict__ nus, int nnu,
                                                 int nmax, double slab, double* __restrict__ out, double* __restrict__ dout) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= npix * nt * nnu) return;
    double glx[16], glw[16];
    for (int q = 0; q < 16; ++q) { glx[q] = GL16_X[q]; glw[q] = GL16_W[q]; }
    tangent_item<KIND>(item, a, thetao, lnL, al, be, npix, N, sp, ns, times, nt, nus, nmax, slab, glx, glw, out, dout);
}

hipError_t launch_tangent(hipStream_t s, int kind, double a, double thetao, double lnL, const double* al, const double* be, int64_t npix, int N,
                          const SplatC* sp, int ns, const double* times, int nt, const double* nus, int nnu, int nmax, double slab,
                          double* out, double* dout) {
    const int64_t items = npix * nt * nnu;
    const dim3 grid(grid_for(items, 64)), block(64);
    switch (kind) {
        case MODEL_THERMAL: hipLaunchKernelGGL(ks_tangent<MODEL_THERMAL>, grid, block, 0, s, a, thetao, lnL, al, be, npix, N, sp, ns, times, nt, nus, nnu, nmax, slab, out, dout); break;
        case MODEL_POWERLAW: hipLaunchKernelGGL(ks_tangent<MODEL_POWERLAW>, grid, block, 0, s, a, thetao, lnL, al, be, npix, N, sp, ns, times, nt, nus, nnu, nmax, slab, out, dout); break;
        case MODEL_KAPPA: hipLaunchKernelGGL(ks_tangent<MODEL_KAPPA>, grid, block, 0, s, a, thetao, lnL, al, be, npix, N, sp, ns, times, nt, nus, nnu, nmax, slab, out, dout); break;
        default: hipLaunchKernelGGL(ks_tangent<MODEL_COMPOSITE>, grid, block, 0, s, a, thetao, lnL, al, be, npix, N, sp, ns, times, nt, nus, nnu, nmax, slab, out, dout); break;
    }
    return hipGetLastError();
}

}  // namespace ksd
