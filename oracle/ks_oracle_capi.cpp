// C entry points of the CPU oracle, loaded by tests/ (through oracle/__init__.py), by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg. TEST INFRASTRUCTURE: the product
// (kerrsplat_amd/libkerrsplat_hip.so) never links or calls this library.
//
// Every entry point works in screen order on host arrays and returns 0 on success. Parallel over
// pixels with OpenMP (one ray per iteration, the reference's KA CPU() backend layout).
#include "ks_splats.hpp"
#include "ks_vis.hpp"
#include "ks_adjoint.hpp"
#include "ks_fluid.hpp"
#include <cstring>
#include <vector>
#include <numeric>
#include <omp.h>

using namespace kso;

namespace {
// the model and its per-splat Bessel table (kept alive by the caller's ModelHolder)
struct ModelHolder { Model m; std::vector<double> bessel; };
template <class T> ModelHolder make_model(int kind, int nrows, int ns, const double* params, const double* hyp, double gamma_max,
                                          const double* knots, const double* knot_times, int nknots) {
    ModelHolder h;
    Model& m = h.m;
    m.kind = kind; m.nrows = nrows; m.ns = ns; m.params = params; m.hyp = hyp; m.gamma_max = gamma_max;
    m.knots = knots; m.knot_times = knot_times; m.nknots = nknots;
    h.bessel.assign(3 * (size_t)std::max(ns, 1), 0.0);
    for (int i = 0; i < ns; ++i) {
        double x = 0.0;
        if (kind == MODEL_THERMAL || kind == MODEL_KNOT) x = 1.0 / std::exp(params[13 + (size_t)nrows * i]);
        else if (kind == MODEL_KAPPA) x = 1.0 / std::exp(params[14 + (size_t)nrows * i]);
        else continue;
        besselk012(x, h.bessel[3 * i], h.bessel[3 * i + 1], h.bessel[3 * i + 2]);
    }
    m.bessel = h.bessel.data();
    return h;
}

// march one ray with the chosen marcher and fold a consumer over its samples
template <class T, class F> void march(const Pixel<T>& p, int N, int M, int marcher, F&& f, T& rt, T& rp) {
    if (marcher == 0) { march_direct(p, N, f); rt = 0.0; rp = 0.0; }
    else march_ray(p, N, M, f, rt, rp);
}

// a ray constant in the spacetime scalar R: the value alone for R = double, else the scalar itself
template <class R, class T> R as_spacetime(const T& x) {
    if constexpr (std::is_same_v<R, double>) return val(x);
    else return R(x);
}

// the polarized transport of one ray (transport.jl:263-290, with the half-orbit counter when nmax >= 0); the
// spin, inclination and length unit in R (double, or the duals of the spacetime tangent with T = R)
template <class T, class R = double> void polarized_ray(const Model& m, R a, R thetao, double alpha, double beta, int N, int M, int marcher,
                                                        double t_obs, double nu, R L, int nmax, double slab, T out[4], T* rt, T* rp) {
    Pixel<T> p = pixel_constants<T>(T(a), T(alpha), T(beta), T(thetao));
    RayCtxT<R> rc{a, R(alpha), R(beta), thetao, as_spacetime<R>(p.eta), as_spacetime<R>(p.lam)};
    RadState<T> st = radstate_zero<T>();
    Winding<T> w = winding_zero<T>();
    const double av = val(a);
    double hor = (1.0 + std::sqrt(1.0 - av * av)) * (1.0 + 1e-3);
    auto f = [&](int k, const Sample<T>& s, T dtau) {
        if (nmax >= 0) {
            if (!(s.ok && s.r > hor)) return;
            wind(w, s.r * kcos(s.th), T(slab));
            if (w.n > nmax) return;
        }
        M4<T> O; T E[4];
        if (sample_step(m, rc, s, dtau, T(t_obs), T(nu), T(L), O, E)) advance(st, O, E);
    };
    T r1, r2;
    march(p, N, M, marcher, f, r1, r2);
    T nh = nu / NU0;
    for (int q = 0; q < 4; ++q) out[q] = st.S[q] * nh * nh * nh;
    if (rt) *rt = r1;
    if (rp) *rp = r2;
}

// the polarized transport of one ray through a fluid model whose elements at a sample are visited by
// elems(rc, s, add) (transport.jl:263-290 with nelements/element of the model): the same screen coefficients,
// step and front-to-back state as the splat models
template <class Elems> void fluid_ray(Elems&& elems, double a, double thetao, double alpha, double beta, int N, int M, int marcher,
                                      double nu, double L, double out[4]) {
    const Pixel<double> p = pixel_constants<double>(a, alpha, beta, thetao);
    const RayCtx rc{a, alpha, beta, thetao, p.eta, p.lam};
    RadState<double> st = radstate_zero<double>();
    const double hor = (1.0 + std::sqrt(1.0 - a * a)) * (1.0 + 1e-3);
    auto f = [&](int, const Sample<double>& s, double dtau) {
        if (!(s.ok && s.r > hor)) return;
        double j4[4] = {0, 0, 0, 0}, a4[4] = {0, 0, 0, 0}, r3[3] = {0, 0, 0};
        bool active = false;
        elems(rc, s, [&](const Coef<double>& cf, const Frame<double>& fr) {
            Screen<double> sc;
            if (!screen_coefficients(cf, fr, nu, sc)) return;
            for (int q = 0; q < 4; ++q) { j4[q] += sc.j[q]; a4[q] += sc.a[q]; }
            for (int q = 0; q < 3; ++q) r3[q] += sc.r[q];
            active = true;
        });
        if (!active) return;
        const double c = std::cos(s.th);
        const double D = L / (nu / NU0) * (s.r * s.r + a * a * c * c) * dtau;
        M4<double> O, Ej;
        step_operator(a4, r3, D, O, Ej);
        double E[4];
        for (int i = 0; i < 4; ++i) E[i] = Ej.m[i][0] * j4[0] + Ej.m[i][1] * j4[1] + Ej.m[i][2] * j4[2] + Ej.m[i][3] * j4[3];
        advance(st, O, E);
    };
    double r1, r2;
    march(p, N, M, marcher, f, r1, r2);
    const double nh = nu / NU0;
    for (int q = 0; q < 4; ++q) out[q] = st.S[q] * nh * nh * nh;
}
}  // namespace

extern "C" {

int kso_num_threads(void) { return omp_get_max_threads(); }
void kso_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// per-pixel constants: out = [nreal, case, tau_total, fo, k, lam, eta, up, um, roots re/im x4 (8)] (17 doubles)
int kso_pixel_constants(double a, double thetao, double alpha, double beta, double* out) {
    Pixel<double> p = pixel_constants<double>(a, alpha, beta, thetao);
    out[0] = p.nreal; out[1] = p.rcase; out[2] = p.tau_total; out[3] = p.fo; out[4] = p.k;
    out[5] = p.lam; out[6] = p.eta; out[7] = p.up; out[8] = p.um;
    for (int i = 0; i < 4; ++i) { out[9 + 2 * i] = p.roots[i].re; out[10 + 2 * i] = p.roots[i].im; }
    return 0;
}

// root case per pixel (screen order): number of real radial roots
int kso_root_cases(double a, double thetao, int npix, const double* alphas, const double* betas, int8_t* numreals) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < npix; ++i) numreals[i] = (int8_t)num_real_roots<double>(a, alphas[i], betas[i], thetao);
    return 0;
}

// stable permutation 4 / 2 / 0 real roots (1-based, as the reference's case_permutation) and the case counts
int kso_case_permutation(int npix, const int8_t* numreals, int64_t* perm, int64_t* counts) {
    std::vector<int64_t> idx(npix);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return numreals[x] > numreals[y]; });
    counts[0] = counts[1] = counts[2] = 0;
    for (int j = 0; j < npix; ++j) {
        perm[j] = idx[j] + 1;
        int n = numreals[idx[j]];
        counts[n == 4 ? 0 : (n == 2 ? 1 : 2)]++;
    }
    return 0;
}

// stored samples (screen order), npix x N row-major per pixel: marcher 0 Direct, 1 Recurrence(M)
int kso_samples(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher,
                double* t, double* r, double* th, double* ph, uint8_t* flags, double* resid_t, double* resid_ph, double* tau_total) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < npix; ++i) {
        Pixel<double> p = pixel_constants<double>(a, alphas[i], betas[i], thetao);
        double rt, rp;
        auto f = [&](int k, const Sample<double>& s, double) {
            size_t o = (size_t)i * N + (k - 1);
            t[o] = s.t; r[o] = s.r; th[o] = s.th; ph[o] = s.ph;
            flags[o] = (s.ok ? SAMPLE_OK : 0) | (s.nur ? SAMPLE_NUR : 0) | (s.nuth ? SAMPLE_NUTH : 0);
        };
        march(p, N, M, marcher, f, rt, rp);
        if (resid_t) resid_t[i] = rt;
        if (resid_ph) resid_ph[i] = rp;
        if (tau_total) tau_total[i] = p.tau_total;
    }
    return 0;
}

// ray_lists (dual_gradient.jl:283-359): hits[i + ns (j + npix f)] = 1 when a sample of ray j that takes a step
// (ok, outside the horizon guard; before the nmax-th half-orbit passage through |z| < slab when nmax >= 0) lies
// inside splat i's support sphere at times[f] (outside_support, polarized.jl:110-120), else 0; screen order.
// Any model whose rows 0-5, 10 and last row are the geometry (thermal, power-law, kappa).
int kso_ray_hits(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher, int nrows, int ns,
                 const double* params, int nt, const double* times, int nmax, double slab, uint8_t* hits) {
    Model m;
    m.nrows = nrows; m.ns = ns; m.params = params;
    const double hor = (1.0 + std::sqrt(1.0 - a * a)) * (1.0 + 1e-3);
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < npix; ++j) {
        Pixel<double> p = pixel_constants<double>(a, alphas[j], betas[j], thetao);
        Winding<double> w = winding_zero<double>();
        auto f = [&](int, const Sample<double>& s, double) {
            if (!(s.ok && s.r > hor)) return;
            if (nmax >= 0) {
                wind(w, s.r * kcos(s.th), slab);
                if (w.n > nmax) return;
            }
            double x, y, z;
            quasi_cartesian_ks(a, s.r, s.th, s.ph, x, y, z);
            for (int fr = 0; fr < nt; ++fr)
                for (int i = 0; i < ns; ++i) {
                    uint8_t& h = hits[(size_t)i + (size_t)ns * ((size_t)j + (size_t)npix * fr)];
                    if (!h && !outside_support(m, i, times[fr] - s.t, x, y, z)) h = 1;
                }
        };
        double rt, rp;
        march(p, N, M, marcher, f, rt, rp);
    }
    return 0;
}

// optically thin Stokes I image (screen order)
int kso_thin_image(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher,
                   int ns, const double* params, double t_obs, double* out) {
    ModelHolder mh = make_model<double>(MODEL_THIN, 14, ns, params, nullptr, 0, nullptr, nullptr, 0);
    const Model& m = mh.m;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < npix; ++i) {
        Pixel<double> p = pixel_constants<double>(a, alphas[i], betas[i], thetao);
        RayCtx rc{a, alphas[i], betas[i], thetao, p.eta, p.lam};
        double acc = 0, rt, rp;
        auto f = [&](int, const Sample<double>& s, double dtau) { acc += thin_sample(m, rc, s, dtau, t_obs); };
        march(p, N, M, marcher, f, rt, rp);
        out[i] = acc;
    }
    return 0;
}

// polarized Stokes image (screen order, npix x 4 observed I, Q, U, V in cgs)
int kso_polarized_image(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher,
                        int kind, int nrows, int ns, const double* params, const double* hyp, double gamma_max,
                        const double* knots, const double* knot_times, int nknots,
                        double t_obs, double nu, double L, int nmax, double slab, double* out, double* resid_t, double* resid_ph) {
    ModelHolder mh = make_model<double>(kind, nrows, ns, params, hyp, gamma_max, knots, knot_times, nknots);
    const Model& m = mh.m;
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < npix; ++i) {
        double rt, rp;
        polarized_ray<double>(m, a, thetao, alphas[i], betas[i], N, M, marcher, t_obs, nu, L, nmax, slab, out + 4 * (size_t)i, &rt, &rp);
        if (resid_t) resid_t[i] = rt;
        if (resid_ph) resid_ph[i] = rp;
    }
    return 0;
}

// spacetime_valgrad / spacetime_jacobian (joint.jl:74-154): the observed Stokes vector of every ray and its
// forward-mode partials with respect to (a, theta_o, ln L), the three duals seeded on the spin, the inclination
// and the log length unit and carried through the pixel constants, the march and the transport.
// out: npix x 4; dout: npix x 4 x 3 (partial fastest).
int kso_spacetime_tangent(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher,
                          int kind, int nrows, int ns, const double* params, const double* hyp, double gamma_max,
                          double t_obs, double nu, double lnL, int nmax, double slab, double* out, double* dout) {
    using D3 = Dual<3, double>;
    ModelHolder mh = make_model<double>(kind, nrows, ns, params, hyp, gamma_max, nullptr, nullptr, 0);
    const Model& m = mh.m;
    if (kind == MODEL_THIN || kind == MODEL_KNOT) return 2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i) {
        const D3 ad = D3::seed(a, 0), td = D3::seed(thetao, 1), Ld = kexp(D3::seed(lnL, 2));
        D3 o[4];
        polarized_ray<D3, D3>(m, ad, td, alphas[i], betas[i], N, M, marcher, t_obs, nu, Ld, nmax, slab, o, nullptr, nullptr);
        for (int q = 0; q < 4; ++q) {
            out[4 * (size_t)i + q] = o[q].v;
            for (int k = 0; k < 3; ++k) dout[12 * (size_t)i + 3 * q + k] = o[q].d[k];
        }
    }
    return 0;
}

// polarized Stokes image of a CompositeModel (transport.jl:298-316): part p has kinds[p], nrows[p] x ns[p]
// parameters params[p], kappa factors hyp[p] (or null) and power-law cutoff gamma_max[p]; the elements
// run over part 0, then part 1, ... Knot parts are not composed here (their knots stay per model).
int kso_polarized_composite(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher,
                            int nparts, const int* kinds, const int* nrows, const int* ns, const double* const* params,
                            const double* const* hyp, const double* gamma_max, const double* const* knots, const double* const* knot_times,
                            const int* nknots, double t_obs, double nu, double L, int nmax, double slab,
                            double* out, double* resid_t, double* resid_ph) {
    if (nparts < 1) return 1;
    std::vector<ModelHolder> parts;
    parts.reserve(nparts);
    for (int p = 0; p < nparts; ++p) {
        if (kinds[p] == MODEL_THIN) return 2;
        if (kinds[p] == MODEL_KNOT && (!knots || !knots[p] || !knot_times[p] || nknots[p] < 1)) return 3;
        const bool kn = knots && knots[p] && knot_times[p] && nknots[p] >= 1;
        parts.push_back(make_model<double>(kinds[p], nrows[p], ns[p], params[p], hyp ? hyp[p] : nullptr, gamma_max[p], kn ? knots[p] : nullptr,
                                           kn ? knot_times[p] : nullptr, kn ? nknots[p] : 0));
    }
    for (int p = 0; p < nparts; ++p) {
        parts[p].m.bessel = parts[p].bessel.data();
        parts[p].m.next = p + 1 < nparts ? &parts[p + 1].m : nullptr;
    }
    const Model& m = parts[0].m;
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < npix; ++i) {
        double rt, rp;
        polarized_ray<double>(m, a, thetao, alphas[i], betas[i], N, M, marcher, t_obs, nu, L, nmax, slab, out + 4 * (size_t)i, &rt, &rp);
        if (resid_t) resid_t[i] = rt;
        if (resid_ph) resid_ph[i] = rp;
    }
    return 0;
}

// algorithmic operation count of the Recurrence(M) march alone over a set of rays (Counted<double>): the samples
// (t, r, theta, phi, momentum signs) of every ray, consumed by nothing
int kso_count_march(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, double* flops,
                    double* special) {
    double f = 0, s = 0;
#pragma omp parallel for reduction(+ : f, s) schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i) {
        op_counts().reset();
        Pixel<Counted> p = pixel_constants<Counted>(Counted(a), Counted(alphas[i]), Counted(betas[i]), Counted(thetao));
        Counted rt, rp;
        march_ray(p, N, M, [](int, const Sample<Counted>&, Counted) {}, rt, rp);
        f += op_counts().flops; s += op_counts().special;
    }
    *flops = f; *special = s;
    return 0;
}

// algorithmic operation count of the polarized forward transport over a set of rays (Counted<double>)
int kso_count_polarized(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M,
                        int kind, int nrows, int ns, const double* params, const double* hyp, double gamma_max,
                        double t_obs, double nu, double L, double* flops, double* special) {
    ModelHolder mh = make_model<Counted>(kind, nrows, ns, params, hyp, gamma_max, nullptr, nullptr, 0);
    const Model& m = mh.m;
    double f = 0, s = 0;
#pragma omp parallel for reduction(+ : f, s) schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i) {
        op_counts().reset();
        Counted out[4];
        polarized_ray<Counted>(m, a, thetao, alphas[i], betas[i], N, M, 1, t_obs, nu, L, -1, 0.0, out, nullptr, nullptr);
        f += op_counts().flops; s += op_counts().special;
    }
    *flops = f; *special = s;
    return 0;
}

// gradient of sum_pixels dstokes . image w.r.t. thermal splat parameters (the dual sweep,
// dual_gradient.jl:225-245): dparams (21 x ns, column-major) accumulates; image (npix x 4) is written
int kso_polarized_gradient(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M,
                           int ns, const double* params, double t_obs, double nu, double L, int nmax, double slab,
                           const double* dstokes, double* dparams, double* image, int analytic_weight) {
    ModelHolder mh = make_model<double>(MODEL_THERMAL, 21, ns, params, nullptr, 0, nullptr, nullptr, 0);
    const Model& m = mh.m;
    int nth = omp_get_max_threads();
    std::vector<std::vector<double>> part(nth, std::vector<double>(21 * (size_t)ns, 0.0));
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i) {
        std::vector<double>& g = part[omp_get_thread_num()];
        if (analytic_weight)
            polarized_gradient_ray<double, true>(m, a, thetao, alphas[i], betas[i], N, M, t_obs, nu, L, nmax, slab, dstokes + 4 * (size_t)i,
                                                 g.data(), image + 4 * (size_t)i);
        else
            polarized_gradient_ray<double, false>(m, a, thetao, alphas[i], betas[i], N, M, t_obs, nu, L, nmax, slab, dstokes + 4 * (size_t)i,
                                                  g.data(), image + 4 * (size_t)i);
    }
    for (int t = 0; t < nth; ++t)
        for (size_t k = 0; k < 21 * (size_t)ns; ++k) dparams[k] += part[t][k];
    return 0;
}

// The dual-sweep gradient for any model: a composite of parts (as kso_polarized_composite; a single population
// is one part; any fluid part may carry knots), every row of every splat differentiated by the generic element
// adjoint. dparams: the parts' nrows x ns blocks in order (accumulates); image (npix x 4) is written.
int kso_model_gradient(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M,
                       int nparts, const int* kinds, const int* nrows, const int* ns, const double* const* params,
                       const double* const* hyp, const double* gamma_max, const double* const* knots, const double* const* knot_times,
                       const int* nknots, double t_obs, double nu, double L, int nmax, double slab,
                       const double* dstokes, double* dparams, double* image) {
    if (nparts < 1) return 1;
    std::vector<ModelHolder> parts;
    parts.reserve(nparts);
    size_t total = 0;
    for (int p = 0; p < nparts; ++p) {
        if (kinds[p] == MODEL_THIN) return 2;
        if (kinds[p] == MODEL_KNOT && (!knots || !knots[p] || !knot_times[p] || nknots[p] < 1)) return 3;
        const bool kn = knots && knots[p] && knot_times[p] && nknots[p] >= 1;
        parts.push_back(make_model<double>(kinds[p], nrows[p], ns[p], params[p], hyp ? hyp[p] : nullptr, gamma_max[p], kn ? knots[p] : nullptr,
                                           kn ? knot_times[p] : nullptr, kn ? nknots[p] : 0));
        total += (size_t)nrows[p] * ns[p];
    }
    for (int p = 0; p < nparts; ++p) {
        parts[p].m.bessel = parts[p].bessel.data();
        parts[p].m.next = p + 1 < nparts ? &parts[p + 1].m : nullptr;
    }
    const Model& m = parts[0].m;
    int nth = omp_get_max_threads();
    std::vector<std::vector<double>> part(nth, std::vector<double>(std::max(total, (size_t)1), 0.0));
#pragma omp parallel for schedule(dynamic, 2)
    for (int i = 0; i < npix; ++i) {
        std::vector<double>& g = part[omp_get_thread_num()];
        polarized_gradient_ray<double, false, true>(m, a, thetao, alphas[i], betas[i], N, M, t_obs, nu, L, nmax, slab, dstokes + 4 * (size_t)i,
                                                    g.data(), image + 4 * (size_t)i);
    }
    for (int t = 0; t < nth; ++t)
        for (size_t k = 0; k < total; ++k) dparams[k] += part[t][k];
    return 0;
}

// the work of the dual sweep per phase on the Counted scalar (with d ln G written out, as the GPU does):
// counts[0..3] = the march, the forward
// steps (splat sums + step operators), the tails, the adjoint sweep, and counts[4] = the part of the
// sweep spent in the per-sample step adjoints (summed over the rays)
int kso_count_gradient(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M,
                       int ns, const double* params, double t_obs, double nu, double L, const double* dstokes,
                       double* counts, double* special) {
    ModelHolder mh = make_model<Counted>(MODEL_THERMAL, 21, ns, params, nullptr, 0, nullptr, nullptr, 0);
    const Model& m = mh.m;
    double c[5] = {0, 0, 0, 0, 0}, sp[5] = {0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : c[:5], sp[:5])
    for (int i = 0; i < npix; ++i) {
        op_counts().reset();
        std::vector<Counted> g(21 * (size_t)ns, Counted(0.0));
        Counted img[4];
        OpCounts ph[5];
        polarized_gradient_ray<Counted, true>(m, a, thetao, alphas[i], betas[i], N, M, t_obs, nu, L, -1, 0.0, dstokes + 4 * (size_t)i, g.data(),
                                              img, ph);
        double pf = 0, ps = 0;
        for (int q = 0; q < 4; ++q) { c[q] += ph[q].flops - pf; sp[q] += ph[q].special - ps; pf = ph[q].flops; ps = ph[q].special; }
        c[4] += ph[4].flops; sp[4] += ph[4].special;
    }
    for (int q = 0; q < 5; ++q) { counts[q] = c[q]; special[q] = sp[q]; }
    return 0;
}

// ---- unit-level entry points (coefficients, step, frames, special functions) ----
int kso_thermal(double ne, double Th, double B, double nu, double th, int dexter_rhoV, int pandya, double* out) {
    Coef<double> c = thermal_synchrotron(ne, Th, B, nu, th, dexter_rhoV != 0, pandya != 0);
    double v[8] = {c.jI, c.jQ, c.jV, c.aI, c.aQ, c.aV, c.rQ, c.rV};
    std::memcpy(out, v, sizeof v);
    return 0;
}
int kso_thermal_pandya(double ne, double Th, double B, double nu, double th, double* out) {
    thermal_pandya(ne, Th, B, nu, th, out[0], out[1], out[2]);
    return 0;
}
double kso_thermal_leung(double ne, double Th, double B, double nu, double th) { return thermal_leung(ne, Th, B, nu, th); }
int kso_powerlaw(double ne, double p, double gmin, double gmax, double B, double nu, double th, int rot, double* out) {
    Coef<double> c = powerlaw_synchrotron(ne, p, gmin, gmax, B, nu, th, rot != 0);
    double v[8] = {c.jI, c.jQ, c.jV, c.aI, c.aQ, c.aV, c.rQ, c.rV};
    std::memcpy(out, v, sizeof v);
    return 0;
}
int kso_kappa(double ne, double kappa, double w, double B, double nu, double th, double hyp, double* out) {
    Coef<double> c = kappa_synchrotron(ne, kappa, w, B, nu, th, hyp);
    double v[8] = {c.jI, c.jQ, c.jV, c.aI, c.aQ, c.aV, c.rQ, c.rV};
    std::memcpy(out, v, sizeof v);
    return 0;
}
double kso_kappa_hyp(double kappa, double w) { return kappa_hypergeometric(kappa, w); }
int kso_besselk01(double x, double* k0, double* k1) { besselk01(x, *k0, *k1); return 0; }
double kso_planck(double nu, double Th) { return planck(nu, Th); }
int kso_step_operator(const double* al, const double* rh, double D, double* O, double* Ej) {
    M4<double> Om, Em;
    step_operator(al, rh, D, Om, Em);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) { O[4 * i + j] = Om.m[i][j]; Ej[4 * i + j] = Em.m[i][j]; }
    return 0;
}
int kso_cap(const double* c8, double* out) {
    Coef<double> c{c8[0], c8[1], c8[2], c8[3], c8[4], c8[5], c8[6], c8[7]};
    Coef<double> r = cap_polarization(c);
    double v[8] = {r.jI, r.jQ, r.jV, r.aI, r.aQ, r.aV, r.rQ, r.rV};
    std::memcpy(out, v, sizeof v);
    return 0;
}
int kso_local_frame(double a, double r, double th, double eta, double lam, int nur, int nuth, double alpha, double beta, double thetao,
                    const double* u, const double* B, double* out) {
    Frame<double> f = local_frame(a, r, th, eta, lam, nur != 0, nuth != 0, alpha, beta, thetao, u, B);
    out[0] = f.g; out[1] = f.cosB; out[2] = f.chi;
    return 0;
}
int kso_elliptic(double phi, double m, double n, double* out) {
    out[0] = ellF(phi, m); out[1] = ellE(phi, m); out[2] = ellPi(n, phi, m);
    Jac<double> j = jacobi(phi, m);
    out[3] = j.sn; out[4] = j.cn; out[5] = j.dn;
    out[6] = jacobi_am_any(phi, m);
    return 0;
}

// visibilities of Stokes images (ks_vis.hpp); images npix x 4 per frame in screen order
int kso_visibilities(int nalpha, int nbeta, int nframes, const double* images, double psize, int nuv, const double* u, const double* v,
                     const double* taper, double* Vre, double* Vim) {
    visibilities(nalpha, nbeta, nframes, images, psize, nuv, u, v, taper, Vre, Vim);
    return 0;
}
int kso_vis_seed(int nalpha, int nbeta, int nframes, double psize, int nuv, const double* u, const double* v, const double* taper,
                 const double* Wre, const double* Wim, double* seed) {
    visibility_seed(nalpha, nbeta, nframes, psize, nuv, u, v, taper, Wre, Wim, seed);
    return 0;
}

}  // extern "C"

extern "C" {
// ipole's example RIAF (test/riaf_model.jl): observed Stokes (I, Q, U, V) per pixel, screen order, out npix x 4
int kso_riaf_image(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher, double nu,
                   double L, int pandya, double* out) {
    const RIAF m = riaf_example(a, pandya != 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i)
        fluid_ray([&](const RayCtx& rc, const Sample<double>& s, auto&& add) {
            Coef<double> cf; Frame<double> fr;
            if (riaf_element(m, rc, s, nu, cf, fr)) add(cf, fr);
        }, a, thetao, alphas[i], betas[i], N, M, marcher, nu, L, out + 4 * (size_t)i);
    return 0;
}

// shells with given fluid-frame coefficients (test_polarized.jl:33-51): slab k is slabs[16 k ..] =
// (r1, r2, jI, jQ, jV, aI, aQ, aV, rQ, rV, u[3], B[3]); observed Stokes per pixel, out npix x 4
int kso_slab_image(double a, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher, double nu,
                   double L, int nslab, const double* slabs, double* out) {
    std::vector<Slab> els((size_t)nslab);
    for (int k = 0; k < nslab; ++k) {
        const double* q = slabs + 16 * (size_t)k;
        els[k] = Slab{q[0], q[1], {q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]}, {q[10], q[11], q[12]}, {q[13], q[14], q[15]}};
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i)
        fluid_ray([&](const RayCtx& rc, const Sample<double>& s, auto&& add) {
            for (const Slab& e : els) {
                Coef<double> cf; Frame<double> fr;
                if (slab_element(e, rc, s, cf, fr)) add(cf, fr);
            }
        }, a, thetao, alphas[i], betas[i], N, M, marcher, nu, L, out + 4 * (size_t)i);
    return 0;
}

// Gold et al. (2020) model `model` (0..4) at its spin (test_gold2020.jl): the observed intensity per pixel
// through the unpolarized transport (transport.jl:36-75), out npix; with polarized the same model through the
// polarized transport as one unpolarized element (test_polarized.jl:24-32), out npix x 4. *spin receives a.
int kso_gold_image(int model, double thetao, int npix, const double* alphas, const double* betas, int N, int M, int marcher, double nu,
                   double L, int polarized, double* out, double* spin) {
    if (model < 0 || model > 4) return 1;
    double a; Gold gm;
    gold_model(model, a, gm);
    if (spin) *spin = a;
    const double nh = nu / NU0;
    const double hor = (1.0 + std::sqrt(1.0 - a * a)) * (1.0 + 1e-3);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npix; ++i) {
        if (polarized) {
            fluid_ray([&](const RayCtx& rc, const Sample<double>& s, auto&& add) {
                double j, al, g;
                gold_coefficients(gm, rc, s, nu, j, al, g);
                add(Coef<double>{j, 0.0, 0.0, al, 0.0, 0.0, 0.0, 0.0}, Frame<double>{g, 0.0, 0.0});
            }, a, thetao, alphas[i], betas[i], N, M, marcher, nu, L, out + 4 * (size_t)i);
            continue;
        }
        const Pixel<double> p = pixel_constants<double>(a, alphas[i], betas[i], thetao);
        const RayCtx rc{a, alphas[i], betas[i], thetao, p.eta, p.lam};
        double tau = 0.0, I = 0.0;
        auto f = [&](int, const Sample<double>& s, double dtau) {
            if (!(s.ok && s.r > hor)) return;
            double j, al, g;
            gold_coefficients(gm, rc, s, nu, j, al, g);
            const double c = std::cos(s.th);
            const double D = L / nh * (s.r * s.r + a * a * c * c) * dtau;
            const double ji = j * g * g / (nh * nh), ai = al * nh / g;
            const double ti = ai * D;
            I += std::exp(-tau) * ji * D * phi1(ti);
            tau += ti;
        };
        double r1, r2;
        march(p, N, M, marcher, f, r1, r2);
        out[i] = I * nh * nh * nh;
    }
    return 0;
}
}  // extern "C"

