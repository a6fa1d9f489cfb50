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
struct PolF {
    const SplatC* __restrict__ sp; const double* __restrict__ knots; const double* __restrict__ ktimes;
    int ns, kind, nknots, nmax;
    double a, rp, rm, thetao_sin, alpha, beta, eta, lam, t_obs, nu, L, slab;
    double P[16], Sv[4];
    double zprev; int wn; bool inside, crossed;

    KS_DEV void init() {
        for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
        for (int i = 0; i < 4; ++i) Sv[i] = 0.0;
        zprev = NAN; wn = 0; inside = false; crossed = false;
    }
    KS_DEV bool wind(double z) {      // step.jl:498-517; returns whether the sample is beyond nmax
        if (isnan(zprev)) { zprev = z; inside = fabs(z) < slab; crossed = false; return wn > nmax; }
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
        return wn > nmax;
    }
    KS_DEV void operator()(int, const Samp& s, double dtau) {
        if (!(s.ok && s.r > rp * (1.0 + 1e-3))) return;
        if (nmax >= 0 && wind(s.r * cos(s.th))) return;
        double x, y, z;
        quasi_cartesian_ks(a, rp, rm, s.r, s.th, s.ph, x, y, z);
        const double t = t_obs - s.t;
        double j4[4] = {0, 0, 0, 0}, a4[4] = {0, 0, 0, 0}, r3[3] = {0, 0, 0};
        bool active = false, geo_ready = false;
        Geo G;
        for (int i = 0; i < ns; ++i) {
            const SplatC& c = sp[i];
            double gexp;
            if (kind == MODEL_KNOT) {
                // centre from the knots (trajectories.jl:197-233): no pattern rotation
                const double* K = knots + 3 * (int64_t)nknots * i;
                double cx, cy, cz;
                if (t <= ktimes[0]) { cx = K[0]; cy = K[1]; cz = K[2]; }
                else if (t >= ktimes[nknots - 1]) { cx = K[3 * (nknots - 1)]; cy = K[3 * (nknots - 1) + 1]; cz = K[3 * (nknots - 1) + 2]; }
                else {
                    int kk = 0;
                    while (kk < nknots - 2 && ktimes[kk + 1] <= t) ++kk;
                    double w = (t - ktimes[kk]) / (ktimes[kk + 1] - ktimes[kk]);
                    cx = (1.0 - w) * K[3 * kk] + w * K[3 * (kk + 1)];
                    cy = (1.0 - w) * K[3 * kk + 1] + w * K[3 * (kk + 1) + 1];
                    cz = (1.0 - w) * K[3 * kk + 2] + w * K[3 * (kk + 1) + 2];
                }
                double d0 = x - cx, d1 = y - cy, d2 = z - cz;
                double u0 = c.R[0] * d0 + c.R[3] * d1 + c.R[6] * d2;
                double u1 = c.R[1] * d0 + c.R[4] * d1 + c.R[7] * d2;
                double u2 = c.R[2] * d0 + c.R[5] * d1 + c.R[8] * d2;
                double q0 = u0 * c.is[0], q1 = u1 * c.is[1], q2 = u2 * c.is[2];
                double tt = (t - c.t0) * c.iw;
                gexp = q0 * q0 + q1 * q1 + q2 * q2 + tt * tt;
            } else {
                double sph, cph;
                if (outside_support(c, t, x, y, z, sph, cph)) continue;
                gexp = gauss_exponent(c, t, x, y, z, sph, cph);
            }
            double G_ = exp(-0.5 * gexp);
            if (!(G_ > WEIGHT_CUTOFF)) continue;
            if (!geo_ready) { make_geo(a, s.r, s.th, eta, lam, s.nur, s.nuth, G); geo_ready = true; }
            Frame fr = local_frame(G, a, alpha, beta, thetao_sin, c.u, c.B, c.Bmag);
            double nuf = nu / fr.g;
            double cb = dmin(dmax(fr.cosB, -1.0), 1.0);
            double sth = sqrt(dmax(1.0 - cb * cb, 0.0)), cth = cb;   // sin, cos of acos(cb)
            {   // the pitch angle through acos, as the reference (safe_acos), for identical sin at the ends
                double thB = acos(cb);
                sincos(thB, &sth, &cth);
            }
            double ne = exp(c.lnA) * G_;
            Coef cf;
            if (kind == MODEL_THERMAL || kind == MODEL_KNOT) cf = thermal(ne, c.p1, c.p2, c.p3, c.Bmag, nuf, sth, cth);
            else if (kind == MODEL_POWERLAW) cf = powerlaw(ne, c.p1, c.p2, c.p3, c.g, c.Bmag, nuf, sth, cth);
            else cf = kappa(ne, c.p1, c.p2, c.p3, c.g, c.Bmag, nuf, sth, cth);
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
        if (!active) return;
        double ct = cos(s.th);
        double Sg = s.r * s.r + a * a * ct * ct;
        double D = L / (nu * (1.0 / NU0)) * Sg * dtau;
        double O[16], E[4];
        transfer_step(j4, a4, r3, D, O, E);
        // S += P E ; P = P O
        for (int i = 0; i < 4; ++i) Sv[i] += P[4 * i] * E[0] + P[4 * i + 1] * E[1] + P[4 * i + 2] * E[2] + P[4 * i + 3] * E[3];
        double Pn[16];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) Pn[4 * i + k] = P[4 * i] * O[k] + P[4 * i + 1] * O[4 + k] + P[4 * i + 2] * O[8 + k] + P[4 * i + 3] * O[12 + k];
        for (int i = 0; i < 16; ++i) P[i] = Pn[i];
    }
};

// one work item per (sorted pixel, time, frequency): item = j + npix (it + nt inu); out[4 item + q] (observed I, Q, U, V)
__global__ void __launch_bounds__(128) ks_march_polarized(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                          const SplatC* __restrict__ sp, int ns, int kind,
                                                          const double* __restrict__ knots, const double* __restrict__ ktimes, int nknots,
                                                          const double* __restrict__ times, int nt, const double* __restrict__ nus, int nnu,
                                                          double L, int nmax, double slab, double* __restrict__ out,
                                                          double* __restrict__ rt, double* __restrict__ rp) {
    int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= npix * nt * nnu) return;
    int64_t j = item % npix;
    int64_t fr = item / npix;
    int it = (int)(fr % nt), inu = (int)(fr / nt);
    Spin S = make_spin(a);
    PixC p = load_pixel(soa, npix, j);
    PolF f;
    f.sp = sp; f.knots = knots; f.ktimes = ktimes; f.ns = ns; f.kind = kind; f.nknots = nknots; f.nmax = nmax;
    f.a = a; f.rp = S.rp; f.rm = S.rm; f.thetao_sin = sin(thetao); f.alpha = p.alpha; f.beta = p.beta; f.eta = p.eta; f.lam = p.lam;
    f.t_obs = times[it]; f.nu = nus[inu]; f.L = L; f.slab = slab;
    f.init();
    double r1, r2;
    march_ray(p, S, thetao, N, M, f, r1, r2);
    double nh = f.nu * (1.0 / NU0);
    double s3 = nh * nh * nh;
    for (int q = 0; q < 4; ++q) out[4 * item + q] = f.Sv[q] * s3;
    if (fr == 0) { rt[j] = r1; rp[j] = r2; }
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
    std::vector<int> perm_host;
};

struct ks_model {
    ks_ctx* ctx = nullptr;
    int kind = KS_MODEL_THERMAL, nrows = 0, ns = 0;
    SplatC* d_splats = nullptr;
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
    *out = c;
    return KS_OK;
}

int ks_ctx_destroy(ks_ctx* c) {
    if (!c) return KS_OK;
    (void)hipSetDevice(c->device);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
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
        c.B[0] = c.Bmag * st * cp; c.B[1] = c.Bmag * st * sp; c.B[2] = c.Bmag * ct;
        c.u[0] = p[rB + 3]; c.u[1] = p[rB + 4]; c.u[2] = p[rB + 5];
        if (kind == KS_MODEL_THERMAL || kind == KS_MODEL_KNOT) {
            double Th = std::exp(p[13]);
            double k0, k1;
            double k2 = host_besselk(1.0 / Th, k0, k1);
            c.p1 = Th;
            c.p2 = k2 > 0.0 ? k0 / k2 : 1.0;
            c.p3 = k2 > 0.0 ? k1 / k2 : 1.0;
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
    if (dalloc(&m->d_splats, h.size())) { delete m; return KS_ERR_HIP; }
    HIPCHK(hipMemcpy(m->d_splats, h.data(), h.size() * sizeof(SplatC), hipMemcpyHostToDevice));
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
    dfree(m->d_splats); dfree(m->d_knots); dfree(m->d_ktimes);
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
    // times and frequencies travel as kernel-argument buffers staged in the output's tail-free scratch
    static thread_local double* d_tn = nullptr; static thread_local size_t d_tn_n = 0;
    if (d_tn_n < (size_t)(nt + nnu)) { dfree(d_tn); if (dalloc(&d_tn, (size_t)(nt + nnu))) return KS_ERR_HIP; d_tn_n = nt + nnu; }
    std::vector<double> tn(times, times + nt); tn.insert(tn.end(), nus, nus + nnu);
    HIPCHK(hipMemcpyAsync(d_tn, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
    int64_t items = c->npix * frames;
    int M = c->marcher == KS_MARCHER_DIRECT ? 64 : c->M;
    hipLaunchKernelGGL(ks_march_polarized, dim3(grid_for(items, 128)), dim3(128), 0, x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M,
                       m->d_splats, m->ns, m->kind, m->d_knots, m->d_ktimes, m->nknots, d_tn, nt, d_tn + nt, nnu, L, nmax, slab, c->out, c->rt, c->rp);
    HIPCHK(hipGetLastError());
    return KS_OK;
}

int ks_render(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
              double* out_stokes, int screen_order) {
    int rc = ks_render_device(c, m, times, nt, nus, nnu, L, nmax, slab);
    if (rc) return rc;
    return ks_copy_stokes(c, (int64_t)nt * nnu, out_stokes, screen_order);
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
