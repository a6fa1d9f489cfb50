// libkerrsplat_hip: MI355X (gfx950) kernels of KerrSplat's fused geodesic march + polarized transfer
// behind the C-ABI declared in include/kerrsplat_hip.h.
//
// Kernels (one work item per ray, case-sorted slots so a wavefront runs one root case):
//   K0  ks_root_case        number of real radial roots per pixel (screen order)
//   --  ks_case_count /     stable 3-bucket partition 4 / 2 / 0 real roots -> perm (sorted -> screen)
//       ks_case_scatter
//   K1  ks_pixel_constants  per-pixel constants, structure of arrays in sorted order
//   K2  ks_march_store      Direct or Recurrence(M) march writing (t~, r, theta, phi, flags) per sample
//       ks_march_thin       fused march + optically thin Stokes I (ThinRenderer)
//       ks_march_polarized  fused march + polarized transport over (pixel, frame, frequency) items
//   --  ks_unsort           sorted -> screen scatter of Stokes images
//   --  ks_vis              visibilities of the images onto (u, v) points
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>

#include "ks_march.hpp"
#include "ks_dtransfer.hpp"
#include "ks_dadjoint.hpp"
#include "../../include/kerrsplat_hip.h"

using namespace ksd;

// ======================================================================================================
// kernels
// ======================================================================================================
__global__ void ks_root_case(double a, double thetao, const double* __restrict__ al, const double* __restrict__ be, int64_t npix,
                             int8_t* __restrict__ numreals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    numreals[i] = (int8_t)root_case_count(a, thetao, al[i], be[i]);
}

// per block: counts of the three cases
constexpr int PART_BLOCK = 256;
__global__ void ks_case_count(const int8_t* __restrict__ numreals, int64_t npix, int* __restrict__ bcount /* 3 x nblocks */) {
    __shared__ int c[3];
    if (threadIdx.x < 3) c[threadIdx.x] = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    if (i < npix) {
        int n = numreals[i];
        atomicAdd(&c[n == 4 ? 0 : (n == 2 ? 1 : 2)], 1);
    }
    __syncthreads();
    if (threadIdx.x < 3) bcount[threadIdx.x * gridDim.x + blockIdx.x] = c[threadIdx.x];
}
// stable scatter: each block writes its pixels of case c at offset[c][block] + (rank within block)
__global__ void ks_case_scatter(const int8_t* __restrict__ numreals, int64_t npix, const int* __restrict__ boff, int* __restrict__ perm) {
    __shared__ int flag[3][PART_BLOCK];
    int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    int cs = -1;
    if (i < npix) { int n = numreals[i]; cs = n == 4 ? 0 : (n == 2 ? 1 : 2); }
    for (int c = 0; c < 3; ++c) flag[c][threadIdx.x] = (cs == c);
    __syncthreads();
    if (cs >= 0) {
        int rank = 0;
        for (int t = 0; t < (int)threadIdx.x; ++t) rank += flag[cs][t];
        perm[boff[cs * gridDim.x + blockIdx.x] + rank] = (int)i;
    }
}

__global__ void ks_pixel_constants(double a, double thetao, const double* __restrict__ al, const double* __restrict__ be,
                                   const int* __restrict__ perm, int64_t npix, double* __restrict__ soa) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    int i = perm[j];
    Spin S = make_spin(a);
    double glx[16], glw[16];
    for (int q = 0; q < 16; ++q) { glx[q] = GL16_X[q]; glw[q] = GL16_W[q]; }
    PixC p = pixel_constants(S, thetao, al[i], be[i], glx, glw);
    store_pixel(soa, npix, j, p);
}

// ---- stored samples ---------------------------------------------------------------------------------
struct StoreF {
    double *t, *r, *th, *ph; uint8_t* fl; int64_t j, npix;
    KS_DEV void operator()(int k, const Samp& s, double) {
        int64_t o = (int64_t)(k - 1) * npix + j;
        t[o] = s.t; r[o] = s.r; th[o] = s.th; ph[o] = s.ph;
        fl[o] = (uint8_t)((s.ok ? 1 : 0) | (s.nur ? 2 : 0) | (s.nuth ? 4 : 0));
    }
};
__global__ void __launch_bounds__(128) ks_march_store(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M, int marcher,
                                                      double* t, double* r, double* th, double* ph, uint8_t* fl, double* rt, double* rp) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    Spin S = make_spin(a);
    PixC p = load_pixel(soa, npix, j);
    StoreF f{t, r, th, ph, fl, j, npix};
    double r1 = 0.0, r2 = 0.0;
    if (marcher == 0) {
        VisC vis = vis_constants(p, a, thetao);
        double dtau = p.tau_total / (double)(N + 1);
        for (int k = 1; k <= N; ++k) f(k, direct_sample(p, S, vis, k * dtau), dtau);
    } else {
        march_ray(p, S, thetao, N, M, f, r1, r2);
    }
    rt[j] = r1; rp[j] = r2;
}

// ---- thin splats (Splats.jl:113-128) --------------------------------------------------------------------
struct ThinF {
    const SplatC* sp; int ns; double a, rp, rm, t_obs, lam, acc;
    KS_DEV void operator()(int, const Samp& s, double dtau) {
        if (!(s.ok && rp * (1.0 + 1e-3) < s.r && s.r < 1e3)) return;
        double x, y, z;
        quasi_cartesian_ks(a, rp, rm, s.r, s.th, s.ph, x, y, z);
        double t = t_obs - s.t;
        double jt = 0.0;
        for (int i = 0; i < ns; ++i) {
            const SplatC& c = sp[i];
            double sph, cph; sincos(c.omega * (t - c.t0), &sph, &cph);
            jt += exp(c.lnA - 0.5 * gauss_exponent(c, t, x, y, z, sph, cph));
        }
        double st, ct; sincos(s.th, &st, &ct);
        double S_ = s.r * s.r + a * a * ct * ct, Dl = s.r * s.r - 2.0 * s.r + a * a;
        double Xi = (s.r * s.r + a * a) * (s.r * s.r + a * a) - a * a * Dl * st * st;
        double g = sqrt(Dl * S_ / Xi) / (1.0 - 2.0 * a * s.r / Xi * lam);
        acc += g * g * jt * S_ * dtau;
    }
};
__global__ void __launch_bounds__(128) ks_march_thin(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                     const SplatC* __restrict__ sp, int ns, double t_obs, double* out) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    Spin S = make_spin(a);
    PixC p = load_pixel(soa, npix, j);
    ThinF f{sp, ns, a, S.rp, S.rm, t_obs, p.lam, 0.0};
    double r1, r2;
    march_ray(p, S, thetao, N, M, f, r1, r2);
    out[j] = f.acc;
}

// ---- polarized transport (transport.jl:263-290) ----------------------------------------------------------
// What a polarized sample needs besides the sample itself.
struct TransCtx {
    const SplatC* __restrict__ sp; const double* __restrict__ knots; const double* __restrict__ ktimes;
    int ns, kind, nknots;
    double a, rp, rm, thetao_sin, alpha, beta, eta, lam, t_obs, nu, L;
};

// the splat sum of one sample (accumulate_elements, transport.jl:254-276): invariant j, alpha, rho in
// screen axes; false when no splat contributes
KS_DEV bool sample_coefficients(const TransCtx& c, const Samp& s, double x, double y, double z, double j4[4], double a4[4], double r3[3]) {
    const double t = c.t_obs - s.t;
    for (int i = 0; i < 4; ++i) { j4[i] = 0.0; a4[i] = 0.0; }
    r3[0] = r3[1] = r3[2] = 0.0;
    bool active = false, geo_ready = false;
    Geo G;
    for (int i = 0; i < c.ns; ++i) {
        const SplatC& sc = c.sp[i];
        double gexp;
        if (c.kind == MODEL_KNOT) {
            // centre from the knots (trajectories.jl:197-233): no pattern rotation
            const double* K = c.knots + 3 * (int64_t)c.nknots * i;
            const int nk = c.nknots;
            double cx, cy, cz;
            if (t <= c.ktimes[0]) { cx = K[0]; cy = K[1]; cz = K[2]; }
            else if (t >= c.ktimes[nk - 1]) { cx = K[3 * (nk - 1)]; cy = K[3 * (nk - 1) + 1]; cz = K[3 * (nk - 1) + 2]; }
            else {
                int kk = 0;
                while (kk < nk - 2 && c.ktimes[kk + 1] <= t) ++kk;
                double w = (t - c.ktimes[kk]) / (c.ktimes[kk + 1] - c.ktimes[kk]);
                cx = (1.0 - w) * K[3 * kk] + w * K[3 * (kk + 1)];
                cy = (1.0 - w) * K[3 * kk + 1] + w * K[3 * (kk + 1) + 1];
                cz = (1.0 - w) * K[3 * kk + 2] + w * K[3 * (kk + 1) + 2];
            }
            double d0 = x - cx, d1 = y - cy, d2 = z - cz;
            double u0 = sc.R[0] * d0 + sc.R[3] * d1 + sc.R[6] * d2;
            double u1 = sc.R[1] * d0 + sc.R[4] * d1 + sc.R[7] * d2;
            double u2 = sc.R[2] * d0 + sc.R[5] * d1 + sc.R[8] * d2;
            double q0 = u0 * sc.is[0], q1 = u1 * sc.is[1], q2 = u2 * sc.is[2];
            double tt = (t - sc.t0) * sc.iw;
            gexp = q0 * q0 + q1 * q1 + q2 * q2 + tt * tt;
        } else {
            double sph, cph;
            if (outside_support(sc, t, x, y, z, sph, cph)) continue;
            gexp = gauss_exponent(sc, t, x, y, z, sph, cph);
        }
        double G_ = exp(-0.5 * gexp);
        if (!(G_ > WEIGHT_CUTOFF)) continue;
        if (!geo_ready) { make_geo(c.a, s.r, s.th, c.eta, c.lam, s.nur, s.nuth, G); geo_ready = true; }
        Frame fr = local_frame(G, c.a, c.alpha, c.beta, c.thetao_sin, sc.u, sc.bh);
        double nuf = c.nu / fr.g;
        double cb = dmin(dmax(fr.cosB, -1.0), 1.0);
        double sth, cth;
        sincos(acos(cb), &sth, &cth);    // the pitch angle through acos, as the reference's safe_acos
        double ne = exp(sc.lnA) * G_;
        Coef cf;
        if (c.kind == MODEL_THERMAL || c.kind == MODEL_KNOT) cf = thermal(ne, sc.p1, sc.p2, sc.p3, sc.Bmag, nuf, sth, cth);
        else if (c.kind == MODEL_POWERLAW) cf = powerlaw(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        else cf = kappa(ne, sc.p1, sc.p2, sc.p3, sc.g, sc.Bmag, nuf, sth, cth);
        if (!(cf.jI > 0.0 || cf.aI > 0.0 || cf.rQ != 0.0 || cf.rV != 0.0)) continue;
        double fe = cap_factor(cf.jI, cf.jQ, cf.jV), fa = cap_factor(cf.aI, cf.aQ, cf.aV);
        double nh = nuf * (1.0 / NU0);
        double i2 = 1.0 / (nh * nh);
        double s2c, c2c; sincos(2.0 * fr.chi, &s2c, &c2c);
        double jQ = fe * cf.jQ * i2, aQ = fa * cf.aQ * nh, rQ = cf.rQ * nh;
        j4[0] += cf.jI * i2; j4[1] += jQ * c2c; j4[2] += jQ * s2c; j4[3] += fe * cf.jV * i2;
        a4[0] += cf.aI * nh; a4[1] += aQ * c2c; a4[2] += aQ * s2c; a4[3] += fa * cf.aV * nh;
        r3[0] += rQ * c2c; r3[1] += rQ * s2c; r3[2] += cf.rV * nh;
        active = true;
    }
    return active;
}

// Delta = (L / nu-hat) Sigma dtau (transport.jl:282-284)
KS_DEV double path_weight(const TransCtx& c, const Samp& s, double dtau) {
    double ct = cos(s.th);
    return c.L / (c.nu * (1.0 / NU0)) * (s.r * s.r + c.a * c.a * ct * ct) * dtau;
}

// A segment's composite transfer map: for its samples k0..k1, front to back, P = O_k0 ... O_k1 and
// Sv = sum_k O_k0 ... O_{k-1} E_k (the RadiativeState of step.jl:342-398 started at the identity).
// Per-sample records of the adjoint (written by the forward segment march when rec != nullptr):
// rec[(f N + k - 1) nitems + item], f = REC_T (emission time t_obs - t~), r, theta, x, y, z, then the
// sample's invariant j (4), alpha (4), rho (3) and Delta; flg[(k - 1) nitems + item] = bit 0 the
// sample takes a step, bit 1 nu_r, bit 2 nu_theta. The adjoint sweep overwrites j, alpha, rho with
// their adjoints j_bar, alpha_bar, rho_bar.
enum RecField { REC_T, REC_R, REC_TH, REC_X, REC_Y, REC_Z, REC_J, REC_A = REC_J + 4, REC_RHO = REC_A + 4, REC_D = REC_RHO + 3, REC_FIELDS };
constexpr int REC_BAR = REC_J;

struct PolSegF {
    TransCtx c;
    int kstop;                 // samples k >= kstop lie past the half-orbit cutoff
    double P[16], Sv[4];
    double* rec; uint8_t* flg; int64_t nitems, item; int N;
    KS_DEV void init() {
        for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
        for (int i = 0; i < 4; ++i) Sv[i] = 0.0;
    }
    KS_DEV void operator()(int k, const Samp& s, double dtau) {
        const int64_t fo = (int64_t)(k - 1) * nitems + item;
        if (!(s.ok && s.r > c.rp * (1.0 + 1e-3)) || k >= kstop) { if (rec) flg[fo] = 0; return; }
        double x, y, z;
        quasi_cartesian_ks(c.a, c.rp, c.rm, s.r, s.th, s.ph, x, y, z);
        double j4[4], a4[4], r3[3];
        if (!sample_coefficients(c, s, x, y, z, j4, a4, r3)) { if (rec) flg[fo] = 0; return; }
        const double D = path_weight(c, s, dtau);
        double O[16], E[4];
        transfer_step(j4, a4, r3, D, O, E);
        // S += P E ; P = P O
        for (int i = 0; i < 4; ++i) Sv[i] += P[4 * i] * E[0] + P[4 * i + 1] * E[1] + P[4 * i + 2] * E[2] + P[4 * i + 3] * E[3];
        double Pn[16];
        for (int i = 0; i < 4; ++i)
            for (int q = 0; q < 4; ++q) Pn[4 * i + q] = P[4 * i] * O[q] + P[4 * i + 1] * O[4 + q] + P[4 * i + 2] * O[8 + q] + P[4 * i + 3] * O[12 + q];
        for (int i = 0; i < 16; ++i) P[i] = Pn[i];
        if (rec) {
            flg[fo] = (uint8_t)(1 | (s.nur ? 2 : 0) | (s.nuth ? 4 : 0));
            const int64_t st = (int64_t)N * nitems;
            double* o = rec + fo;
            o[REC_T * st] = c.t_obs - s.t; o[REC_R * st] = s.r; o[REC_TH * st] = s.th;
            o[REC_X * st] = x; o[REC_Y * st] = y; o[REC_Z * st] = z;
            for (int i = 0; i < 4; ++i) { o[(REC_J + i) * st] = j4[i]; o[(REC_A + i) * st] = a4[i]; }
            for (int i = 0; i < 3; ++i) o[(REC_RHO + i) * st] = r3[i];
            o[REC_D * st] = D;
        }
    }
};

// The half-orbit counter (WindingState, step.jl:498-517) over the valid samples outside the horizon
// guard; kstop = the first sample beyond the nmax-th passage (N + 1 when there is none).
struct WindF {
    double rp, slab; int nmax, kstop;
    double zprev; int wn; bool inside, crossed;
    KS_DEV void operator()(int k, const Samp& s, double) {
        if (kstop <= k || !(s.ok && s.r > rp * (1.0 + 1e-3))) return;
        double z = s.r * cos(s.th);
        if (isnan(zprev)) { zprev = z; inside = fabs(z) < slab; crossed = false; }
        else {
            bool inside_now = fabs(z) < slab;
            bool crossing = zprev * z < 0.0;
            if (inside || inside_now) {
                crossed = crossed || crossing;
                if (inside && !inside_now) { if (crossed) wn += 1; crossed = false; }
                else if (!inside && inside_now) { crossed = crossing; }
                inside = inside_now;
            } else if (crossing) {
                wn += 1;
            }
            zprev = z;
        }
        if (wn > nmax) kstop = k;
    }
};
__global__ void __launch_bounds__(128) ks_wind_cutoff(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                      int nmax, double slab, int* __restrict__ kstop) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    Spin S = make_spin(a);
    PixC p = load_pixel(soa, npix, j);
    WindF f{S.rp, slab, nmax, N + 1, NAN, 0, false, false};
    double r1, r2;
    march_ray(p, S, thetao, N, M, f, r1, r2);
    kstop[j] = f.kstop;
}

KS_DEV TransCtx make_ctx(const Spin& S, const PixC& p, double thetao_sin, const SplatC* sp, int ns, int kind, const double* knots,
                         const double* ktimes, int nknots, double t_obs, double nu, double L) {
    TransCtx c;
    c.sp = sp; c.knots = knots; c.ktimes = ktimes; c.ns = ns; c.kind = kind; c.nknots = nknots;
    c.a = S.a; c.rp = S.rp; c.rm = S.rm; c.thetao_sin = thetao_sin; c.alpha = p.alpha; c.beta = p.beta; c.eta = p.eta; c.lam = p.lam;
    c.t_obs = t_obs; c.nu = nu; c.L = L;
    return c;
}

// Work items are (segment m, item) with item = j + npix (it + nt inu) fastest, so a wavefront holds
// consecutive case-sorted pixels of one segment index. Segment m of pixel j covers samples
// m L + 1 .. min((m + 1) L, N), L = segment_length of the pixel; items with m past the pixel's last
// segment exit. Output: seg[(20 m + f) nitems + item], f = P (16, row-major) then Sv (4); the
// segment's anchor residuals res[2 (m npix + j) + {0, 1}] from frame 0.
constexpr int SEG_FIELDS = 20;
__global__ void __launch_bounds__(128) ks_seg_polarized(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                        const SplatC* __restrict__ sp, int ns, int kind,
                                                        const double* __restrict__ knots, const double* __restrict__ ktimes, int nknots,
                                                        const double* __restrict__ times, int nt, const double* __restrict__ nus, int nnu,
                                                        double L, const int* __restrict__ kstop, int nseg_max,
                                                        double* __restrict__ seg, double* __restrict__ res,
                                                        double* __restrict__ rec, uint8_t* __restrict__ flg) {
    const int64_t nitems = npix * nt * nnu;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nitems * nseg_max) return;
    const int64_t item = tid % nitems;
    const int m = (int)(tid / nitems);
    const int64_t j = item % npix;
    const int64_t fr = item / npix;
    const int it = (int)(fr % nt), inu = (int)(fr / nt);
    const PixC p = load_pixel(soa, npix, j);
    const int Lseg = segment_length(M, 1.0 - p.k);
    const int kbeg = m * Lseg + 1;
    if (kbeg > N) return;
    const int kend = kbeg + Lseg - 1 < N ? kbeg + Lseg - 1 : N;
    const Spin S = make_spin(a);
    PolSegF f;
    f.c = make_ctx(S, p, sin(thetao), sp, ns, kind, knots, ktimes, nknots, times[it], nus[inu], L);
    f.kstop = kstop ? kstop[j] : N + 1;
    f.rec = rec; f.flg = flg; f.nitems = nitems; f.item = item; f.N = N;
    f.init();
    double r1 = 0.0, r2 = 0.0;
    const double dtau = p.tau_total / (double)(N + 1);
    // a segment that starts past the outer horizon holds no valid sample and no valid anchor
    if (kbeg * dtau <= p.tau_h) march_range(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
    else if (rec) for (int k = kbeg; k <= kend; ++k) flg[(int64_t)(k - 1) * nitems + item] = 0;
    double* o = seg + (int64_t)SEG_FIELDS * m * nitems + item;
    for (int q = 0; q < 16; ++q) o[q * nitems] = f.P[q];
    for (int q = 0; q < 4; ++q) o[(16 + q) * nitems] = f.Sv[q];
    if (fr == 0) { res[2 * (m * npix + j)] = r1; res[2 * (m * npix + j) + 1] = r2; }
}

// Compose the segments of an item front to back: S = sum_m P_0 ... P_{m-1} S_m, the observed Stokes
// vector nu-hat^3 S (observed_stokes, transport.jl:311-316); per-pixel residuals from frame 0.
__global__ void ks_seg_combine(const double* __restrict__ soa, int64_t npix, int N, int M, int nt, const double* __restrict__ nus, int nnu,
                               const double* __restrict__ seg, const double* __restrict__ res, double* __restrict__ out,
                               double* __restrict__ rt, double* __restrict__ rp) {
    const int64_t nitems = npix * nt * nnu;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const int64_t j = item % npix;
    const int64_t fr = item / npix;
    const int inu = (int)(fr / nt);
    const double k = soa[(int64_t)PF_K * npix + j];
    const int Lseg = segment_length(M, 1.0 - k);
    const int nseg = (N + Lseg - 1) / Lseg;
    double P[16], Sv[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
    double r1 = 0.0, r2 = 0.0;
    for (int m = 0; m < nseg; ++m) {
        const double* g = seg + (int64_t)SEG_FIELDS * m * nitems + item;
        double Pm[16], Sm[4];
        for (int q = 0; q < 16; ++q) Pm[q] = g[q * nitems];
        for (int q = 0; q < 4; ++q) Sm[q] = g[(16 + q) * nitems];
        for (int i = 0; i < 4; ++i) Sv[i] += P[4 * i] * Sm[0] + P[4 * i + 1] * Sm[1] + P[4 * i + 2] * Sm[2] + P[4 * i + 3] * Sm[3];
        double Pn[16];
        for (int i = 0; i < 4; ++i)
            for (int q = 0; q < 4; ++q) Pn[4 * i + q] = P[4 * i] * Pm[q] + P[4 * i + 1] * Pm[4 + q] + P[4 * i + 2] * Pm[8 + q] + P[4 * i + 3] * Pm[12 + q];
        for (int i = 0; i < 16; ++i) P[i] = Pn[i];
        if (fr == 0) { r1 = dmax(r1, res[2 * (m * npix + j)]); r2 = dmax(r2, res[2 * (m * npix + j) + 1]); }
    }
    const double nh = nus[inu] * (1.0 / NU0);
    const double s3 = nh * nh * nh;
    for (int q = 0; q < 4; ++q) out[4 * item + q] = Sv[q] * s3;
    if (fr == 0) { rt[j] = r1; rp[j] = r2; }
}

// ---- the dual-sweep adjoint (dual_gradient.jl:125-249) ---------------------------------------------------------
// Per item: from the segments' transfer maps (P_m, S_m), the tail behind each segment
// T_out[m] = S_{m+1} + P_{m+1} T_out[m+1] (T_out[last] = 0), the image nu-hat^3 T_in[0], and the adjoint
// in front of each segment w_in[m] = (P_0 ... P_{m-1})^T w_1 with w_1 = nu-hat^3 dstokes
// (dual_gradient.jl:162). wT[(8 m + q) nitems + item] = w_in, wT[(8 m + 4 + q) nitems + item] = T_out.
__global__ void ks_adj_combine(const double* __restrict__ soa, int64_t npix, int N, int M, int nt, const double* __restrict__ nus, int nnu,
                               const double* __restrict__ seg, const double* __restrict__ dstokes, double* __restrict__ wT,
                               double* __restrict__ out) {
    const int64_t nitems = npix * nt * nnu;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const int64_t j = item % npix;
    const int inu = (int)((item / npix) / nt);
    const int Lseg = segment_length(M, 1.0 - soa[(int64_t)PF_K * npix + j]);
    const int nseg = (N + Lseg - 1) / Lseg;
    const double nh = nus[inu] * (1.0 / NU0);
    const double s3 = nh * nh * nh;
    double T[4] = {0, 0, 0, 0};
    for (int m = nseg - 1; m >= 0; --m) {
        const double* g = seg + (int64_t)SEG_FIELDS * m * nitems + item;
        for (int q = 0; q < 4; ++q) wT[(8 * (int64_t)m + 4 + q) * nitems + item] = T[q];
        double Tn[4];
        for (int i = 0; i < 4; ++i) {
            double v = g[(16 + i) * nitems];
            for (int q = 0; q < 4; ++q) v += g[(4 * i + q) * nitems] * T[q];
            Tn[i] = v;
        }
        for (int i = 0; i < 4; ++i) T[i] = Tn[i];
    }
    for (int q = 0; q < 4; ++q) out[4 * item + q] = T[q] * s3;
    double w[4];
    for (int q = 0; q < 4; ++q) w[q] = dstokes[4 * item + q] * s3;
    for (int m = 0; m < nseg; ++m) {
        const double* g = seg + (int64_t)SEG_FIELDS * m * nitems + item;
        for (int q = 0; q < 4; ++q) wT[(8 * (int64_t)m + q) * nitems + item] = w[q];
        double wn[4];
        for (int c = 0; c < 4; ++c) {
            double v = 0.0;
            for (int i = 0; i < 4; ++i) v += g[(4 * i + c) * nitems] * w[i];
            wn[c] = v;
        }
        for (int q = 0; q < 4; ++q) w[q] = wn[q];
    }
}

// Per (segment, item): the tails T_{k+1} of the segment's samples back to front from T_out[m], then
// the adjoint front to back from w_in[m]: per sample the adjoints of its j, alpha, rho (written over
// them in the record) and w <- O^T w (polarized_dual_kernel!'s ray loop, dual_gradient.jl:151-178).
__global__ void __launch_bounds__(128) ks_adj_segment(const double* __restrict__ soa, int64_t npix, int N, int M, int64_t nitems, int nseg_max,
                                                      const double* __restrict__ wT, double* __restrict__ rec, const uint8_t* __restrict__ flg,
                                                      double* __restrict__ tails) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nitems * nseg_max) return;
    const int64_t item = tid % nitems;
    const int m = (int)(tid / nitems);
    const int64_t j = item % npix;
    const int Lseg = segment_length(M, 1.0 - soa[(int64_t)PF_K * npix + j]);
    const int kbeg = m * Lseg + 1;
    if (kbeg > N) return;
    const int kend = kbeg + Lseg - 1 < N ? kbeg + Lseg - 1 : N;
    const int64_t st = (int64_t)N * nitems;
    double T[4], w[4];
    for (int q = 0; q < 4; ++q) { w[q] = wT[(8 * (int64_t)m + q) * nitems + item]; T[q] = wT[(8 * (int64_t)m + 4 + q) * nitems + item]; }
    for (int k = kend; k >= kbeg; --k) {
        const int64_t fo = (int64_t)(k - 1) * nitems + item;
        if (!(flg[fo] & 1)) continue;
        const double* o = rec + fo;
        double j4[4], a4[4], r3[3];
        for (int i = 0; i < 4; ++i) { j4[i] = o[(REC_J + i) * st]; a4[i] = o[(REC_A + i) * st]; }
        for (int i = 0; i < 3; ++i) r3[i] = o[(REC_RHO + i) * st];
        for (int q = 0; q < 4; ++q) tails[(int64_t)q * st + fo] = T[q];
        double Tn[4];
        step_apply(j4, a4, r3, o[REC_D * st], T, Tn);
        for (int q = 0; q < 4; ++q) T[q] = Tn[q];
    }
    for (int k = kbeg; k <= kend; ++k) {
        const int64_t fo = (int64_t)(k - 1) * nitems + item;
        if (!(flg[fo] & 1)) continue;
        double* o = rec + fo;
        double j4[4], a4[4], r3[3], R[4];
        for (int i = 0; i < 4; ++i) { j4[i] = o[(REC_J + i) * st]; a4[i] = o[(REC_A + i) * st]; R[i] = tails[(int64_t)i * st + fo]; }
        for (int i = 0; i < 3; ++i) r3[i] = o[(REC_RHO + i) * st];
        double Otw[4], jb[4], ab[4], rb[3];
        sample_adjoint(j4, a4, r3, o[REC_D * st], w, R, Otw, jb, ab, rb);
        for (int i = 0; i < 4; ++i) { o[(REC_BAR + i) * st] = jb[i]; o[(REC_BAR + 4 + i) * st] = ab[i]; }
        for (int i = 0; i < 3; ++i) o[(REC_BAR + 8 + i) * st] = rb[i];
        for (int q = 0; q < 4; ++q) w[q] = Otw[q];
    }
}

// Per (64 consecutive items, splat): every step-taking sample of the items through the splat's
// element adjoint, 21 partial sums per lane, summed over the wavefront; part[(ib ns + i) 21 + r].
constexpr int ADJ_ROWS = 21;
__global__ void __launch_bounds__(64) ks_adj_splats(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N,
                                                    const SplatC* __restrict__ sp, const SplatAdj* __restrict__ sadj, int ns,
                                                    const double* __restrict__ nus, int nt, int64_t nitems,
                                                    const double* __restrict__ rec, const uint8_t* __restrict__ flg, double* __restrict__ part) {
    const int i = (int)(blockIdx.x % ns);
    const int64_t ib = blockIdx.x / ns;
    const int64_t item = ib * 64 + threadIdx.x;
    double grad[ADJ_ROWS];
    for (int r = 0; r < ADJ_ROWS; ++r) grad[r] = 0.0;
    if (item < nitems) {
        const int64_t j = item % npix;
        const int inu = (int)((item / npix) / nt);
        const double nu = nus[inu];
        const SplatC c = sp[i];
        const SplatAdj sa = sadj[i];
        const Spin S = make_spin(a);
        const double alpha = soa[(int64_t)PF_ALPHA * npix + j], beta = soa[(int64_t)PF_BETA * npix + j];
        const double eta = soa[(int64_t)PF_ETA * npix + j], lam = soa[(int64_t)PF_LAM * npix + j];
        const double tos = sin(thetao);
        const int64_t st = (int64_t)N * nitems;
        for (int k = 1; k <= N; ++k) {
            const int64_t fo = (int64_t)(k - 1) * nitems + item;
            const uint8_t fl = flg[fo];
            if (!(fl & 1)) continue;
            const double* o = rec + fo;
            const double t = o[REC_T * st], x = o[REC_X * st], y = o[REC_Y * st], z = o[REC_Z * st];
            double dl[13];
            const double Gw = weight_dlog(c, sa, t, x, y, z, dl);
            if (!(Gw > 0.0)) continue;
            double bars[11];
            for (int q = 0; q < 11; ++q) bars[q] = o[(REC_BAR + q) * st];
            Geo G;
            make_geo(a, o[REC_R * st], o[REC_TH * st], eta, lam, (fl & 2) != 0, (fl & 4) != 0, G);
            element_adjoint_thermal(c, sa, G, S.a, alpha, beta, tos, nu, bars, Gw, dl, grad);
        }
    }
    // wavefront sum (blockDim = 64 = one wavefront)
    for (int r = 0; r < ADJ_ROWS; ++r) {
        double v = grad[r];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        grad[r] = v;
    }
    if (threadIdx.x == 0) {
        double* o = part + ((int64_t)ib * ns + i) * ADJ_ROWS;
        for (int r = 0; r < ADJ_ROWS; ++r) o[r] = grad[r];
    }
}

// dparams[r + 21 i] = sum over item blocks in ascending order (deterministic)
__global__ void ks_adj_reduce(const double* __restrict__ part, int64_t nib, int ns, double* __restrict__ dparams) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ns * ADJ_ROWS) return;
    const int i = (int)(t / ADJ_ROWS), r = (int)(t % ADJ_ROWS);
    double s = 0.0;
    for (int64_t b = 0; b < nib; ++b) s += part[(b * ns + i) * ADJ_ROWS + r];
    dparams[t] = s;
}

// sorted -> screen: dst[4 (perm[j] + npix f) + q] = src[4 (j + npix f) + q]
__global__ void ks_unsort4(const double* __restrict__ src, const int* __restrict__ perm, int64_t npix, int64_t nframes, double* __restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix * nframes) return;
    int64_t j = i % npix, f = i / npix;
    int64_t d = perm[j] + npix * f;
    for (int q = 0; q < 4; ++q) dst[4 * d + q] = src[4 * i + q];
}
__global__ void ks_unsort1(const double* __restrict__ src, const int* __restrict__ perm, int64_t npix, double* __restrict__ dst) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    dst[perm[j]] = src[j];
}

// ---- visibilities: one work item per (uv point, frame); images in sorted order with perm ---------------------------
__global__ void ks_vis_direct(const double* __restrict__ img, const int* __restrict__ perm, int64_t npix, int nalpha, int nbeta, int nframes,
                              double psize, double scale, const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ tap,
                              int nuv, double* __restrict__ V /* nframes x nuv x 4 complex (re, im) */) {
    int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (int64_t)nuv * nframes) return;
    int k = (int)(item % nuv);
    int64_t f = item / nuv;
    double uk = u[k], vk = v[k];
    double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
    const double* I = img + 4 * npix * f;
    for (int64_t j = 0; j < npix; ++j) {
        int pix = perm[j];
        int ia = pix % nalpha, ib = pix / nalpha;
        double l = -((ia + 1) - (nalpha + 1) * 0.5) * psize, m = ((ib + 1) - (nbeta + 1) * 0.5) * psize;
        double s, c; sincos(2.0 * PI * (uk * l + vk * m), &s, &c);
        for (int q = 0; q < 4; ++q) { double x = I[4 * j + q] * scale; re[q] += x * c; im[q] += x * s; }
    }
    double tk = tap ? tap[k] : 1.0;
    for (int q = 0; q < 4; ++q) { V[2 * (4 * item + q)] = re[q] * tk; V[2 * (4 * item + q) + 1] = im[q] * tk; }
}

// ======================================================================================================
// host side: the C-ABI
// ======================================================================================================
namespace {
thread_local std::string g_err;
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap; va_start(ap, fmt); vsnprintf(buf, sizeof buf, fmt, ap); va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail(KS_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

template <class T> int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return 0;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(KS_ERR_HIP, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}
template <class T> void dfree(T*& p) { if (p) { (void)hipFree(p); p = nullptr; } }

unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
}  // namespace

struct ks_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evk0 = nullptr, evk1 = nullptr;   // around the last dominant-kernel launch
    bool kernel_timed = false;
};

struct ks_cache {
    ks_ctx* ctx = nullptr;
    int64_t npix = 0;
    int N = 0;
    int store = 0;
    int nalpha = 0, nbeta = 1;
    double spin = NAN, thetao = NAN;
    int marcher = KS_MARCHER_DIRECT, M = 64;
    int generation = 0;
    int64_t counts[3] = {0, 0, 0};
    double *al = nullptr, *be = nullptr;
    int8_t* numreals = nullptr;
    int* perm = nullptr;
    int* bcount = nullptr; int nblocks = 0;
    double* soa = nullptr;
    double *st = nullptr, *sr = nullptr, *sth = nullptr, *sph = nullptr; uint8_t* sfl = nullptr;
    double *rt = nullptr, *rp = nullptr;
    double* out = nullptr; int64_t out_frames = 0;    // sorted-order Stokes images
    double* out_screen = nullptr;
    double* img1 = nullptr;                            // thin image (sorted)
    double *vu = nullptr, *vv = nullptr, *vt = nullptr, *vV = nullptr;   // visibility buffers
    int64_t vis_cap = 0, visV_cap = 0;
    double* seg = nullptr; int64_t seg_cap = 0;        // per-segment transfer maps
    double* dtn = nullptr; int64_t dtn_cap = 0;        // observation times then frequencies (device)
    std::vector<double> htn;                           // their host staging copy (outlives the async copy)
    double* segres = nullptr;                          // per-segment anchor residuals (2 x nseg_max x npix)
    int* kstop = nullptr;                              // half-orbit cutoff per pixel
    // adjoint workspace (grow-only): records, flags, tails, per-segment (w_in, T_out), seeds, partial sums
    double* arec = nullptr; uint8_t* aflg = nullptr; double* atail = nullptr; double* awT = nullptr; double* ads = nullptr;
    double* apart = nullptr; double* adpar = nullptr;
    int64_t a_items = 0, a_part = 0, a_dpar = 0;
    std::vector<int> perm_host;
};

struct ks_model {
    ks_ctx* ctx = nullptr;
    int kind = KS_MODEL_THERMAL, nrows = 0, ns = 0;
    SplatC* d_splats = nullptr;
    SplatAdj* d_adj = nullptr;                         // thermal splats: constants of the adjoint
    double* d_knots = nullptr; double* d_ktimes = nullptr; int nknots = 0;
};

extern "C" {

const char* ks_last_error(void) { return g_err.c_str(); }

int ks_ctx_create(int device, ks_ctx** out) {
    if (!out) return fail(KS_ERR_ARG, "ks_ctx_create: null output pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(KS_ERR_ARG, "device %d outside 0:%d", device, n - 1);
    HIPCHK(hipSetDevice(device));
    ks_ctx* c = new ks_ctx;
    c->device = device;
    HIPCHK(hipStreamCreate(&c->stream));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    HIPCHK(hipEventCreate(&c->evk0));
    HIPCHK(hipEventCreate(&c->evk1));
    *out = c;
    return KS_OK;
}

int ks_ctx_destroy(ks_ctx* c) {
    if (!c) return KS_OK;
    (void)hipSetDevice(c->device);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evk0) (void)hipEventDestroy(c->evk0);
    if (c->evk1) (void)hipEventDestroy(c->evk1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return KS_OK;
}

int ks_ctx_sync(ks_ctx* c) {
    if (!c) return fail(KS_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return KS_OK;
}

void* ks_ctx_stream(ks_ctx* c) { return c ? (void*)c->stream : nullptr; }

int ks_timing_begin(ks_ctx* c) {
    if (!c) return fail(KS_ERR_ARG, "null context");
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    return KS_OK;
}
int ks_timing_end(ks_ctx* c, float* ms) {
    if (!c || !ms) return fail(KS_ERR_ARG, "null argument");
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
    return KS_OK;
}
int ks_last_kernel_ms(ks_ctx* c, float* ms) {
    if (!c || !ms) return fail(KS_ERR_ARG, "null argument");
    if (!c->kernel_timed) return fail(KS_ERR_STATE, "no timed kernel launched on this context yet");
    HIPCHK(hipEventSynchronize(c->evk1));
    HIPCHK(hipEventElapsedTime(ms, c->evk0, c->evk1));
    return KS_OK;
}

int ks_cache_create(ks_ctx* ctx, int64_t npix, int N, int store_samples, ks_cache** out) {
    if (!ctx || !out) return fail(KS_ERR_ARG, "ks_cache_create: null argument");
    if (N < 1) return fail(KS_ERR_ARG, "need at least one sample per ray");
    if (npix < 1) return fail(KS_ERR_ARG, "need at least one pixel");
    if (npix > ((int64_t)1 << 30)) return fail(KS_ERR_ARG, "npix %lld above the supported 2^30", (long long)npix);
    HIPCHK(hipSetDevice(ctx->device));
    ks_cache* c = new ks_cache;
    c->ctx = ctx; c->npix = npix; c->N = N; c->store = store_samples ? 1 : 0;
    c->nalpha = (int)npix; c->nbeta = 1;
    c->nblocks = (int)grid_for(npix, PART_BLOCK);
    int rc = 0;
    rc |= dalloc(&c->al, npix); rc |= dalloc(&c->be, npix); rc |= dalloc(&c->numreals, npix); rc |= dalloc(&c->perm, npix);
    rc |= dalloc(&c->bcount, 3 * (size_t)c->nblocks);
    rc |= dalloc(&c->soa, (size_t)PF_COUNT * npix);
    rc |= dalloc(&c->rt, npix); rc |= dalloc(&c->rp, npix);
    if (c->store) {
        size_t ns = (size_t)npix * N;
        rc |= dalloc(&c->st, ns); rc |= dalloc(&c->sr, ns); rc |= dalloc(&c->sth, ns); rc |= dalloc(&c->sph, ns); rc |= dalloc(&c->sfl, ns);
    }
    if (rc) { int code = rc; ks_cache_destroy(c); return code ? KS_ERR_HIP : KS_ERR_HIP; }
    HIPCHK(hipMemsetAsync(c->rt, 0, npix * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(c->rp, 0, npix * sizeof(double), ctx->stream));
    c->perm_host.resize(npix);
    for (int64_t i = 0; i < npix; ++i) c->perm_host[i] = (int)i;
    *out = c;
    return KS_OK;
}

int ks_cache_destroy(ks_cache* c) {
    if (!c) return KS_OK;
    (void)hipSetDevice(c->ctx->device);
    dfree(c->al); dfree(c->be); dfree(c->numreals); dfree(c->perm); dfree(c->bcount); dfree(c->soa);
    dfree(c->st); dfree(c->sr); dfree(c->sth); dfree(c->sph); dfree(c->sfl); dfree(c->rt); dfree(c->rp);
    dfree(c->out); dfree(c->out_screen); dfree(c->img1);
    dfree(c->vu); dfree(c->vv); dfree(c->vt); dfree(c->vV);
    dfree(c->seg); dfree(c->segres); dfree(c->kstop); dfree(c->dtn);
    dfree(c->arec); dfree(c->aflg); dfree(c->atail); dfree(c->awT); dfree(c->ads); dfree(c->apart); dfree(c->adpar);
    delete c;
    return KS_OK;
}

int ks_cache_info(const ks_cache* c, int64_t* npix, int* N, int* generation, int64_t counts[3]) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (npix) *npix = c->npix;
    if (N) *N = c->N;
    if (generation) *generation = c->generation;
    if (counts) for (int i = 0; i < 3; ++i) counts[i] = c->counts[i];
    return KS_OK;
}

int ks_regenerate(ks_cache* c, const double* alpha, const double* beta, int nalpha, int nbeta, double a, double thetao, int marcher, int M) {
    if (!c || !alpha || !beta) return fail(KS_ERR_ARG, "ks_regenerate: null argument");
    if ((int64_t)nalpha * nbeta != c->npix) return fail(KS_ERR_DIM, "camera has %lld pixels, cache was built for %lld", (long long)nalpha * nbeta, (long long)c->npix);
    if (marcher != KS_MARCHER_DIRECT && marcher != KS_MARCHER_RECURRENCE && marcher != KS_MARCHER_FUSED)
        return fail(KS_ERR_ARG, "unknown marcher %d", marcher);
    if (marcher != KS_MARCHER_FUSED && !c->store)
        return fail(KS_ERR_ARG, "this cache was built without sample storage; use marcher = Fused(M) and fused_march!");
    if (M < 1) return fail(KS_ERR_ARG, "re-anchoring interval must be positive");
    if (!(a > 0.0 && a < 1.0)) return fail(KS_ERR_ARG, "spin a = %g outside (0, 1)", a);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    hipStream_t s = x->stream;
    const int64_t n = c->npix;
    c->nalpha = nalpha; c->nbeta = nbeta;
    HIPCHK(hipMemcpyAsync(c->al, alpha, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->be, beta, n * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(ks_root_case, dim3(grid_for(n, 256)), dim3(256), 0, s, a, thetao, c->al, c->be, n, c->numreals);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(ks_case_count, dim3(c->nblocks), dim3(PART_BLOCK), 0, s, c->numreals, n, c->bcount);
    HIPCHK(hipGetLastError());
    // block offsets (the three case sizes and per-block exclusive prefixes) on the host: 3 x nblocks ints
    std::vector<int> bc(3 * (size_t)c->nblocks), off(3 * (size_t)c->nblocks);
    HIPCHK(hipMemcpyAsync(bc.data(), c->bcount, bc.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int64_t base = 0;
    for (int cs = 0; cs < 3; ++cs) {
        int64_t tot = 0;
        for (int b = 0; b < c->nblocks; ++b) { off[cs * c->nblocks + b] = (int)(base + tot); tot += bc[cs * c->nblocks + b]; }
        c->counts[cs] = tot;
        base += tot;
    }
    HIPCHK(hipMemcpyAsync(c->bcount, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(ks_case_scatter, dim3(c->nblocks), dim3(PART_BLOCK), 0, s, c->numreals, n, c->bcount, c->perm);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(ks_pixel_constants, dim3(grid_for(n, 128)), dim3(128), 0, s, a, thetao, c->al, c->be, c->perm, n, c->soa);
    HIPCHK(hipGetLastError());
    c->spin = a; c->thetao = thetao; c->marcher = marcher; c->M = M;
    if (marcher != KS_MARCHER_FUSED) {
        hipLaunchKernelGGL(ks_march_store, dim3(grid_for(n, 128)), dim3(128), 0, s, a, thetao, c->soa, n, c->N, M,
                           marcher == KS_MARCHER_DIRECT ? 0 : 1, c->st, c->sr, c->sth, c->sph, c->sfl, c->rt, c->rp);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(c->perm_host.data(), c->perm, n * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->generation += 1;
    return KS_OK;
}

int ks_get_perm(const ks_cache* c, int64_t* perm, int64_t counts[3]) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (perm) for (int64_t j = 0; j < c->npix; ++j) perm[j] = (int64_t)c->perm_host[j] + 1;
    if (counts) for (int i = 0; i < 3; ++i) counts[i] = c->counts[i];
    return KS_OK;
}

int ks_get_residuals(const ks_cache* c, double* rt, double* rp) {
    if (!c || !rt || !rp) return fail(KS_ERR_ARG, "null argument");
    HIPCHK(hipMemcpy(rt, c->rt, c->npix * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rp, c->rp, c->npix * sizeof(double), hipMemcpyDeviceToHost));
    return KS_OK;
}

int ks_copy_consts(const ks_cache* c, double* soa) {
    if (!c || !soa) return fail(KS_ERR_ARG, "null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    HIPCHK(hipMemcpy(soa, c->soa, (size_t)PF_COUNT * c->npix * sizeof(double), hipMemcpyDeviceToHost));
    return KS_OK;
}
int ks_num_const_fields(void) { return PF_COUNT; }

int ks_copy_samples(const ks_cache* c, double* t, double* r, double* th, double* ph, uint8_t* flags) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (!c->store) return fail(KS_ERR_STATE, "the cache holds no stored samples: build it with store_samples = true and a storing marcher");
    if (c->generation < 1 || c->marcher == KS_MARCHER_FUSED) return fail(KS_ERR_STATE, "no samples: regenerate! with Direct() or Recurrence(M) first");
    size_t n = (size_t)c->npix * c->N;
    if (t) HIPCHK(hipMemcpy(t, c->st, n * sizeof(double), hipMemcpyDeviceToHost));
    if (r) HIPCHK(hipMemcpy(r, c->sr, n * sizeof(double), hipMemcpyDeviceToHost));
    if (th) HIPCHK(hipMemcpy(th, c->sth, n * sizeof(double), hipMemcpyDeviceToHost));
    if (ph) HIPCHK(hipMemcpy(ph, c->sph, n * sizeof(double), hipMemcpyDeviceToHost));
    if (flags) HIPCHK(hipMemcpy(flags, c->sfl, n, hipMemcpyDeviceToHost));
    return KS_OK;
}

// ---- models ----
static double host_besselk(double x, double& k0, double& k1) {
    // K_nu(x) = int_0^inf exp(-x cosh s) cosh(nu s) ds, trapezoidal rule (exponentially convergent)
    const double h = 0.125;
    double s0 = 0.5 * std::exp(-x), s1 = s0;
    for (int i = 1; i < 4000; ++i) {
        double s = h * i;
        double e = std::exp(-x * std::cosh(s));
        s0 += e; s1 += e * std::cosh(s);
        if (e < 1e-40 * s0) break;
    }
    k0 = h * s0; k1 = h * s1;
    return k0 + 2.0 * k1 / x;
}
static double host_hyp2f1_neg(double a, double b, double c, double z) {
    double w = z / (z - 1.0), bb = c - b, term = 1.0, sum = 1.0;
    for (int n = 0; n < 200000; ++n) {
        term *= (a + n) * (bb + n) / ((c + n) * (n + 1.0)) * w;
        sum += term;
        if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
    }
    return std::pow(1.0 - z, -a) * sum;
}

int ks_model_create(ks_ctx* ctx, int kind, int nrows, int ns, const double* params, double gamma_max,
                    const double* knots, const double* knot_times, int nknots, ks_model** out) {
    if (!ctx || !out) return fail(KS_ERR_ARG, "ks_model_create: null argument");
    int want = kind == KS_MODEL_THIN ? 14 : (kind == KS_MODEL_THERMAL || kind == KS_MODEL_KNOT) ? 21 : (kind == KS_MODEL_POWERLAW || kind == KS_MODEL_KAPPA) ? 22 : -1;
    if (want < 0) return fail(KS_ERR_ARG, "unknown model kind %d", kind);
    if (nrows != want) return fail(KS_ERR_DIM, "model kind %d takes %d parameter rows, got %d", kind, want, nrows);
    if (ns < 0 || (ns > 0 && !params)) return fail(KS_ERR_ARG, "bad splat count or parameters");
    if (kind == KS_MODEL_KNOT && (nknots < 1 || !knots || !knot_times)) return fail(KS_ERR_ARG, "knot splats need knots and knot times");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<SplatC> h(ns > 0 ? ns : 1);
    std::vector<SplatAdj> ha(ns > 0 ? ns : 1);
    const double SR2 = -2.0 * std::log(1e-6);
    for (int i = 0; i < ns; ++i) {
        const double* p = params + (size_t)nrows * i;
        SplatC& c = h[i];
        std::memset(&c, 0, sizeof c);
        c.x = p[0]; c.y = p[1]; c.z = p[2];
        double q1 = p[6], q2 = p[7], q3 = p[8], q4 = p[9];
        double nq = std::sqrt(q1 * q1 + q2 * q2 + q3 * q3 + q4 * q4);
        double w = q1 / nq, x = q2 / nq, y = q3 / nq, z = q4 / nq;
        double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                       2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                       2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
        std::memcpy(c.R, R, sizeof R);
        for (int a_ = 0; a_ < 3; ++a_) c.is[a_] = std::exp(-p[3 + a_]);
        c.t0 = p[10]; c.iw = std::exp(-p[11]);
        c.omega = p[nrows - 1];
        double smax = std::max(std::max(p[3], p[4]), p[5]);
        c.rsup2 = SR2 * std::exp(2.0 * smax);
        c.lnA = p[12];
        if (kind == KS_MODEL_THIN) continue;
        int rB = (kind == KS_MODEL_THERMAL || kind == KS_MODEL_KNOT) ? 14 : 15;
        c.Bmag = std::exp(p[rB]);
        double st = std::sin(p[rB + 1]), ct = std::cos(p[rB + 1]), sp = std::sin(p[rB + 2]), cp = std::cos(p[rB + 2]);
        c.bh[0] = st * cp; c.bh[1] = st * sp; c.bh[2] = ct;
        SplatAdj& sa = ha[i];
        std::memset(&sa, 0, sizeof sa);
        sa.qn[0] = w; sa.qn[1] = x; sa.qn[2] = y; sa.qn[3] = z; sa.iqn = 1.0 / nq;
        sa.sthB = st; sa.cthB = ct; sa.sphB = sp; sa.cphB = cp;
        c.u[0] = p[rB + 3]; c.u[1] = p[rB + 4]; c.u[2] = p[rB + 5];
        if (kind == KS_MODEL_THERMAL || kind == KS_MODEL_KNOT) {
            double Th = std::exp(p[13]);
            double k0, k1;
            const double xb = 1.0 / Th;
            double k2 = host_besselk(xb, k0, k1);
            c.p1 = Th;
            c.p2 = k2 > 0.0 ? k0 / k2 : 1.0;
            c.p3 = k2 > 0.0 ? k1 / k2 : 1.0;
            // d/d ln Theta = -x d/dx at x = 1/Theta, with K0' = -K1, K1' = -K0 - K1/x, K2 = K0 + 2 K1/x
            if (k2 > 0.0) {
                const double dk1 = -k0 - k1 / xb, dk2 = -k1 - 2.0 * k0 / xb - 4.0 * k1 / (xb * xb);
                sa.dk0k2 = -xb * (-k1 * k2 - k0 * dk2) / (k2 * k2);
                sa.dk1k2 = -xb * (dk1 * k2 - k1 * dk2) / (k2 * k2);
            }
        } else if (kind == KS_MODEL_POWERLAW) {
            double pp = p[13];
            c.p1 = pp; c.p2 = std::exp(p[14]); c.p3 = gamma_max;
            c.g[0] = std::tgamma((3 * pp - 1) / 12.0) * std::tgamma((3 * pp + 19) / 12.0);
            c.g[1] = std::tgamma((3 * pp + 2) / 12.0) * std::tgamma((3 * pp + 22) / 12.0);
        } else {
            double kap = p[13], wk = std::exp(p[14]);
            c.p1 = kap; c.p2 = wk; c.p3 = host_hyp2f1_neg(kap - 1.0 / 3.0, kap + 1.0, kap + 2.0 / 3.0, -kap * wk);
            c.g[0] = std::tgamma(kap - 4.0 / 3.0); c.g[1] = std::tgamma(kap - 2.0);
            c.g[2] = std::tgamma(kap / 4.0 - 1.0 / 3.0); c.g[3] = std::tgamma(kap / 4.0 + 4.0 / 3.0);
            c.g[4] = std::tgamma(2.0 + kap / 2.0);
            double k0, k1;
            double k2 = host_besselk(1.0 / wk, k0, k1);
            c.g[5] = k0 / (k2 + 1e-40);
        }
    }
    ks_model* m = new ks_model;
    m->ctx = ctx; m->kind = kind; m->nrows = nrows; m->ns = ns;
    if (dalloc(&m->d_splats, h.size()) || dalloc(&m->d_adj, ha.size())) { ks_model_destroy(m); return KS_ERR_HIP; }
    HIPCHK(hipMemcpy(m->d_splats, h.data(), h.size() * sizeof(SplatC), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->d_adj, ha.data(), ha.size() * sizeof(SplatAdj), hipMemcpyHostToDevice));
    if (kind == KS_MODEL_KNOT) {
        m->nknots = nknots;
        if (dalloc(&m->d_knots, (size_t)3 * nknots * ns) || dalloc(&m->d_ktimes, (size_t)nknots)) { ks_model_destroy(m); return KS_ERR_HIP; }
        HIPCHK(hipMemcpy(m->d_knots, knots, (size_t)3 * nknots * ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_ktimes, knot_times, (size_t)nknots * sizeof(double), hipMemcpyHostToDevice));
    }
    *out = m;
    return KS_OK;
}

int ks_model_destroy(ks_model* m) {
    if (!m) return KS_OK;
    (void)hipSetDevice(m->ctx->device);
    dfree(m->d_splats); dfree(m->d_adj); dfree(m->d_knots); dfree(m->d_ktimes);
    delete m;
    return KS_OK;
}

static int ensure_out(ks_cache* c, int64_t frames) {
    if (c->out_frames >= frames && c->out) return KS_OK;
    dfree(c->out); dfree(c->out_screen);
    if (dalloc(&c->out, (size_t)4 * c->npix * frames) || dalloc(&c->out_screen, (size_t)4 * c->npix * frames)) return KS_ERR_HIP;
    c->out_frames = frames;
    return KS_OK;
}

// the observation times then the frequencies, on the device (the host copy is kept by the cache until
// the next call, after the stream has moved past the copy)
static const double* stage_times(ks_cache* c, const double* times, int nt, const double* nus, int nnu) {
    ks_ctx* x = c->ctx;
    if (c->dtn_cap < nt + nnu) {
        dfree(c->dtn);
        if (dalloc(&c->dtn, (size_t)(nt + nnu))) return nullptr;
        c->dtn_cap = nt + nnu;
    }
    if (!c->htn.empty() && hipStreamSynchronize(x->stream) != hipSuccess) { fail(KS_ERR_HIP, "stream synchronize failed"); return nullptr; }
    c->htn.assign(times, times + nt);
    c->htn.insert(c->htn.end(), nus, nus + nnu);
    hipError_t e = hipMemcpyAsync(c->dtn, c->htn.data(), c->htn.size() * sizeof(double), hipMemcpyHostToDevice, x->stream);
    if (e != hipSuccess) { fail(KS_ERR_HIP, "hipMemcpyAsync(times): %s", hipGetErrorString(e)); return nullptr; }
    return c->dtn;
}

int ks_render_device(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab) {
    if (!c || !m || !times || !nus) return fail(KS_ERR_ARG, "ks_render: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (m->kind == KS_MODEL_THIN) return fail(KS_ERR_ARG, "thin splats render through ks_thin_image");
    if (nt < 1 || nnu < 1) return fail(KS_ERR_ARG, "need at least one time and one frequency");
    if (!(L > 0.0)) return fail(KS_ERR_ARG, "length unit L must be positive");
    for (int i = 0; i < nnu; ++i) if (!(nus[i] > 0.0)) return fail(KS_ERR_ARG, "frequencies must be positive");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    int64_t frames = (int64_t)nt * nnu;
    if (ensure_out(c, frames)) return KS_ERR_HIP;
    const double* d_tn = stage_times(c, times, nt, nus, nnu);
    if (!d_tn) return KS_ERR_HIP;
    int64_t items = c->npix * frames;
    int M = c->marcher == KS_MARCHER_DIRECT ? 64 : c->M;
    const int nseg = max_segments(c->N);
    if (c->seg_cap < (int64_t)SEG_FIELDS * nseg * items) {
        dfree(c->seg);
        if (dalloc(&c->seg, (size_t)SEG_FIELDS * nseg * items)) return KS_ERR_HIP;
        c->seg_cap = (int64_t)SEG_FIELDS * nseg * items;
    }
    if (!c->segres) { if (dalloc(&c->segres, (size_t)2 * nseg * c->npix)) return KS_ERR_HIP; }
    const int* kst = nullptr;
    if (nmax >= 0) {
        if (!c->kstop) { if (dalloc(&c->kstop, (size_t)c->npix)) return KS_ERR_HIP; }
        hipLaunchKernelGGL(ks_wind_cutoff, dim3(grid_for(c->npix, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                           nmax, slab, c->kstop);
        HIPCHK(hipGetLastError());
        kst = c->kstop;
    }
    HIPCHK(hipEventRecord(x->evk0, x->stream));
    hipLaunchKernelGGL(ks_seg_polarized, dim3(grid_for(items * nseg, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                       m->d_splats, m->ns, m->kind, m->d_knots, m->d_ktimes, m->nknots, d_tn, nt, d_tn + nt, nnu, L, kst, nseg, c->seg, c->segres,
                       nullptr, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(x->evk1, x->stream));
    x->kernel_timed = true;
    hipLaunchKernelGGL(ks_seg_combine, dim3(grid_for(items, 128)), dim3(128), 0, x->stream, c->soa, c->npix, c->N, M, nt, d_tn + nt, nnu,
                       c->seg, c->segres, c->out, c->rt, c->rp);
    HIPCHK(hipGetLastError());
    return KS_OK;
}

int ks_render(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
              double* out_stokes, int screen_order) {
    int rc = ks_render_device(c, m, times, nt, nus, nnu, L, nmax, slab);
    if (rc) return rc;
    return ks_copy_stokes(c, (int64_t)nt * nnu, out_stokes, screen_order);
}

// ---- the dual-sweep gradient (polarized_gradient!, polarized.jl:473-480; dual_gradient.jl:180-249) ----
static int ensure_adj(ks_cache* c, int64_t items, int ns) {
    const int nseg = max_segments(c->N);
    if (c->a_items < items) {
        dfree(c->arec); dfree(c->aflg); dfree(c->atail); dfree(c->awT); dfree(c->ads);
        c->a_items = 0;
        const size_t per = (size_t)c->N * items;
        if (dalloc(&c->arec, (size_t)REC_FIELDS * per) || dalloc(&c->aflg, per) || dalloc(&c->atail, 4 * per) ||
            dalloc(&c->awT, (size_t)8 * nseg * items) || dalloc(&c->ads, (size_t)4 * items))
            return KS_ERR_HIP;
        c->a_items = items;
    }
    const int64_t nib = (items + 63) / 64;
    if (c->a_part < nib * ns * ADJ_ROWS) {
        dfree(c->apart);
        if (dalloc(&c->apart, (size_t)nib * ns * ADJ_ROWS)) return KS_ERR_HIP;
        c->a_part = nib * ns * ADJ_ROWS;
    }
    if (c->a_dpar < (int64_t)ns * ADJ_ROWS) {
        dfree(c->adpar);
        if (dalloc(&c->adpar, (size_t)ns * ADJ_ROWS)) return KS_ERR_HIP;
        c->a_dpar = (int64_t)ns * ADJ_ROWS;
    }
    return KS_OK;
}

int ks_polarized_gradient(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
                          const double* dstokes, int screen_order, double* dparams, double* image) {
    if (!c || !m || !times || !nus || !dstokes || !dparams) return fail(KS_ERR_ARG, "ks_polarized_gradient: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (m->kind != KS_MODEL_THERMAL) return fail(KS_ERR_ARG, "the dual sweep takes thermal splats (PolarizedSplats)");
    if (nt < 1 || nnu < 1) return fail(KS_ERR_ARG, "need at least one time and one frequency");
    if (!(L > 0.0)) return fail(KS_ERR_ARG, "length unit L must be positive");
    for (int i = 0; i < nnu; ++i) if (!(nus[i] > 0.0)) return fail(KS_ERR_ARG, "frequencies must be positive");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int64_t frames = (int64_t)nt * nnu;
    const int64_t items = c->npix * frames;
    const int ns = m->ns;
    if (ensure_out(c, frames) || ensure_adj(c, items, ns > 0 ? ns : 1)) return KS_ERR_HIP;
    const double* d_tn = stage_times(c, times, nt, nus, nnu);
    if (!d_tn) return KS_ERR_HIP;
    const int M = c->marcher == KS_MARCHER_DIRECT ? 64 : c->M;
    const int nseg = max_segments(c->N);
    if (c->seg_cap < (int64_t)SEG_FIELDS * nseg * items) {
        dfree(c->seg);
        if (dalloc(&c->seg, (size_t)SEG_FIELDS * nseg * items)) return KS_ERR_HIP;
        c->seg_cap = (int64_t)SEG_FIELDS * nseg * items;
    }
    if (!c->segres) { if (dalloc(&c->segres, (size_t)2 * nseg * c->npix)) return KS_ERR_HIP; }
    // seeds in sorted order: dstokes[4 (j + npix f) + q]
    std::vector<double> ds((size_t)4 * items);
    for (int64_t f = 0; f < frames; ++f)
        for (int64_t j = 0; j < c->npix; ++j) {
            const int64_t src = screen_order ? c->perm_host[j] + c->npix * f : j + c->npix * f;
            for (int q = 0; q < 4; ++q) ds[4 * (j + c->npix * f) + q] = dstokes[4 * src + q];
        }
    HIPCHK(hipMemcpyAsync(c->ads, ds.data(), ds.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
    const int* kst = nullptr;
    if (nmax >= 0) {
        if (!c->kstop) { if (dalloc(&c->kstop, (size_t)c->npix)) return KS_ERR_HIP; }
        hipLaunchKernelGGL(ks_wind_cutoff, dim3(grid_for(c->npix, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                           nmax, slab, c->kstop);
        HIPCHK(hipGetLastError());
        kst = c->kstop;
    }
    // forward: segment maps plus the per-sample records
    hipLaunchKernelGGL(ks_seg_polarized, dim3(grid_for(items * nseg, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                       m->d_splats, ns, m->kind, m->d_knots, m->d_ktimes, m->nknots, d_tn, nt, d_tn + nt, nnu, L, kst, nseg, c->seg, c->segres,
                       c->arec, c->aflg);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(ks_adj_combine, dim3(grid_for(items, 128)), dim3(128), 0, x->stream, c->soa, c->npix, c->N, M, nt, d_tn + nt, nnu,
                       c->seg, c->ads, c->awT, c->out);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(ks_adj_segment, dim3(grid_for(items * nseg, 128)), dim3(128), 0, x->stream, c->soa, c->npix, c->N, M, items, nseg,
                       c->awT, c->arec, c->aflg, c->atail);
    HIPCHK(hipGetLastError());
    const int64_t nib = (items + 63) / 64;
    if (ns > 0) {
        HIPCHK(hipEventRecord(x->evk0, x->stream));
        hipLaunchKernelGGL(ks_adj_splats, dim3((unsigned)(nib * ns)), dim3(64), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N,
                           m->d_splats, m->d_adj, ns, d_tn + nt, nt, items, c->arec, c->aflg, c->apart);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(x->evk1, x->stream));
        x->kernel_timed = true;
        hipLaunchKernelGGL(ks_adj_reduce, dim3(grid_for((int64_t)ns * ADJ_ROWS, 64)), dim3(64), 0, x->stream, c->apart, nib, ns, c->adpar);
        HIPCHK(hipGetLastError());
    }
    std::vector<double> g((size_t)ns * ADJ_ROWS);
    if (ns > 0) HIPCHK(hipMemcpyAsync(g.data(), c->adpar, g.size() * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    if (image) {
        const double* src = c->out;
        if (screen_order) {
            hipLaunchKernelGGL(ks_unsort4, dim3(grid_for(c->npix * frames, 256)), dim3(256), 0, x->stream, c->out, c->perm, c->npix, frames, c->out_screen);
            HIPCHK(hipGetLastError());
            src = c->out_screen;
        }
        HIPCHK(hipMemcpyAsync(image, src, (size_t)4 * items * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    for (size_t i = 0; i < g.size(); ++i) dparams[i] += g[i];
    return KS_OK;
}

int ks_copy_stokes(ks_cache* c, int64_t frames, double* out_stokes, int screen_order) {
    if (!c || !out_stokes) return fail(KS_ERR_ARG, "null argument");
    if (frames > c->out_frames) return fail(KS_ERR_DIM, "only %lld frames rendered", (long long)c->out_frames);
    ks_ctx* x = c->ctx;
    size_t n = (size_t)4 * c->npix * frames;
    const double* src = c->out;
    if (screen_order) {
        hipLaunchKernelGGL(ks_unsort4, dim3(grid_for(c->npix * frames, 256)), dim3(256), 0, x->stream, c->out, c->perm, c->npix, frames, c->out_screen);
        HIPCHK(hipGetLastError());
        src = c->out_screen;
    }
    HIPCHK(hipMemcpyAsync(out_stokes, src, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

const double* ks_stokes_device_ptr(const ks_cache* c) { return c ? c->out : nullptr; }

int ks_thin_image(ks_cache* c, const ks_model* m, double t_obs, double* out, int screen_order) {
    if (!c || !m || !out) return fail(KS_ERR_ARG, "ks_thin_image: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (m->kind != KS_MODEL_THIN) return fail(KS_ERR_ARG, "ks_thin_image takes a thin-splat model");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    if (!c->img1) { if (dalloc(&c->img1, (size_t)2 * c->npix)) return KS_ERR_HIP; }
    int M = c->marcher == KS_MARCHER_DIRECT ? 64 : c->M;
    hipLaunchKernelGGL(ks_march_thin, dim3(grid_for(c->npix, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                       m->d_splats, m->ns, t_obs, c->img1);
    HIPCHK(hipGetLastError());
    const double* src = c->img1;
    if (screen_order) {
        hipLaunchKernelGGL(ks_unsort1, dim3(grid_for(c->npix, 256)), dim3(256), 0, x->stream, c->img1, c->perm, c->npix, c->img1 + c->npix);
        HIPCHK(hipGetLastError());
        src = c->img1 + c->npix;
    }
    HIPCHK(hipMemcpyAsync(out, src, c->npix * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

// visibilities of the last rendered frames (images in the cache, sorted order), EHT sign convention
int ks_visibilities(ks_cache* c, int64_t frames, double psize, const double* u, const double* v, const double* taper, int nuv, double* V) {
    if (!c || !u || !v || !V) return fail(KS_ERR_ARG, "ks_visibilities: null argument");
    if (frames > c->out_frames || frames < 1) return fail(KS_ERR_DIM, "only %lld frames rendered", (long long)c->out_frames);
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if ((int64_t)c->nalpha * c->nbeta != c->npix) return fail(KS_ERR_STATE, "the cache's screen shape is not set");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    if (c->vis_cap < nuv) {
        dfree(c->vu); dfree(c->vv); dfree(c->vt);
        if (dalloc(&c->vu, nuv) || dalloc(&c->vv, nuv) || dalloc(&c->vt, nuv)) return KS_ERR_HIP;
        c->vis_cap = nuv;
    }
    if (c->visV_cap < (int64_t)8 * nuv * frames) {
        dfree(c->vV);
        if (dalloc(&c->vV, (size_t)8 * nuv * frames)) return KS_ERR_HIP;
        c->visV_cap = (int64_t)8 * nuv * frames;
    }
    HIPCHK(hipMemcpyAsync(c->vu, u, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->vv, v, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    if (taper) HIPCHK(hipMemcpyAsync(c->vt, taper, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    double scale = psize * psize / 1e-23;
    hipLaunchKernelGGL(ks_vis_direct, dim3(grid_for((int64_t)nuv * frames, 64)), dim3(64), 0, x->stream, c->out, c->perm, c->npix, c->nalpha, c->nbeta,
                       (int)frames, psize, scale, c->vu, c->vv, taper ? c->vt : nullptr, nuv, c->vV);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(V, c->vV, (size_t)8 * nuv * frames * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

}  // extern "C"
