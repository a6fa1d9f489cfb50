// Shared layouts of the device buffers and the host launchers of every kernel of libkerrsplat_hip.
//
// The kernels live in four translation units compiled separately for gfx950:
//   ks_k_geo.hip       K0 root cases, the stable case partition, K1 pixel constants, the stored-sample
//                      march (anchor segments in parallel), the half-orbit cutoff, sorted -> screen
//   ks_k_transfer.hip  the polarized transfer from stored samples (chunks) or fused with the march
//                      (anchor segments), thin images, the composition of the segment maps
//   ks_k_adjoint.hip   the dual-sweep adjoint over stored samples
//   ks_k_vis.hip       visibilities
// ks_lib.hip holds the C-ABI and calls the launchers declared here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "ks_march.hpp"
#include "ks_dtransfer.hpp"
#include "ks_dadjoint.hpp"

namespace ksd {

constexpr int PART_BLOCK = 256;    // block of the case partition
constexpr int CHUNK = 32;          // samples per work item of the transfer from stored samples
constexpr int SEG_FIELDS = 20;     // a segment map: P (16, row-major), then Sv (4)
// the populations an element-adjoint launch meets (a bit mask)
constexpr int POP_THERMAL = 1, POP_POWERLAW = 2, POP_KNOTS = 4, POP_KAPPA = 8;
constexpr int POP_FLUIDS = POP_THERMAL | POP_POWERLAW | POP_KAPPA;
constexpr int ADJ_ROWS = 22;       // gradient rows per splat (21 thermal and knot, 22 power law and kappa)
constexpr int ADJ_SLICE = 32;      // samples per work unit of the element adjoints
constexpr int CAND_MAX = 1024;     // splats a tile's candidate list holds (more splats: no per-tile cull)
constexpr int CULL_MIN_NS = 64;    // the transfer culls per tile from this many splats on (below, the per-sample test is cheaper)
constexpr int UB_FIELDS = 8;       // a unit's region: centre (3), radius, emission times [min, max], any sample, pad
// the minimum and maximum over a wavefront (butterfly: every lane ends with the wavefront's value)
KS_DEV double wave_min(double v) { for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64)); return v; }
KS_DEV double wave_max(double v) { for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64)); return v; }
// A region's bounding sphere from per-lane boxes lo/hi (x, y, z) and emission times (index 3), reduced over
// the wavefront; false when no lane saw a sample.
KS_DEV bool wave_region(double lo[4], double hi[4], double& px, double& py, double& pz, double& R, double& tmin, double& tmax) {
    for (int q = 0; q < 4; ++q) { lo[q] = wave_min(lo[q]); hi[q] = wave_max(hi[q]); }
    if (!(hi[0] >= lo[0])) return false;
    px = 0.5 * (lo[0] + hi[0]); py = 0.5 * (lo[1] + hi[1]); pz = 0.5 * (lo[2] + hi[2]);
    const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    R = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
    tmin = lo[3]; tmax = hi[3];
    return true;
}
// element-adjoint work units (64 items x ADJ_SLICE samples) of nitems items of N samples
inline int64_t adj_units(int64_t nitems, int N) { return ((nitems + 63) / 64) * ((N + ADJ_SLICE - 1) / ADJ_SLICE); }
// Minimum waves per SIMD requested for the heavy kernels (the register budget the compiler may use:
// 512 / waves VGPRs per lane), per precision; tuned by measurement on MI355X.
#ifndef KS_WAVES_MARCH
#define KS_WAVES_MARCH 1
#endif
#ifndef KS_WAVES_TRANSFER64
#define KS_WAVES_TRANSFER64 2
#endif
#ifndef KS_WAVES_TRANSFER32
#define KS_WAVES_TRANSFER32 4
#endif
#ifndef KS_WAVES_ADJSTEP64
#define KS_WAVES_ADJSTEP64 1
#endif
#ifndef KS_WAVES_ADJSTEP32
#define KS_WAVES_ADJSTEP32 2
#endif
#ifndef KS_WAVES_ADJELEM64
#define KS_WAVES_ADJELEM64 1
#endif
#ifndef KS_WAVES_ADJELEM32
#define KS_WAVES_ADJELEM32 2
#endif
template <class S> struct Waves;
template <> struct Waves<double> { static constexpr int transfer = KS_WAVES_TRANSFER64, adjstep = KS_WAVES_ADJSTEP64, adjelem = KS_WAVES_ADJELEM64; };
template <> struct Waves<f32> { static constexpr int transfer = KS_WAVES_TRANSFER32, adjstep = KS_WAVES_ADJSTEP32, adjelem = KS_WAVES_ADJELEM32; };

// FP32 transfer and adjoint: samples whose Faraday depth |rho| Delta exceeds this take the FP64 step
// (the step's polynomial in K' cancels strongly at large Faraday depth)
constexpr double FARADAY_FP64 = 10.0;

// Stored samples, sample-major o = (k - 1) npix + j: t~, r, theta, phi and flags (bit 0 ok, 1 nu_r,
// 2 nu_theta) as the reference stores them (direct_march.jl:14-17), plus the sample's quasi-Cartesian
// Kerr-Schild position (coordinates.jl:22-31), which every render of the samples needs.
struct Samples {
    double *t, *r, *th, *ph, *x, *y, *z;
    uint8_t* fl;
    const int* tmap;      // the sorted pixels in wavefront order (8 x 8 screen tiles), or null for sorted order
    const float4* p4 = nullptr;   // (x, y, z, t~) of every sample in FP32 for the FP32 element adjoints, or null
};
// the item of work slot t (t = j + npix f): the pixel at slot t % npix of the wavefront order, same frame
KS_DEV int64_t slot_item(const Samples& smp, int64_t t, int64_t npix) {
    return smp.tmap ? (int64_t)smp.tmap[t % npix] + npix * (t / npix) : t;
}

// Adjoint records of a step-taking sample: rec[(f N + k - 1) nitems + item], f = the invariant j (4),
// alpha (4), rho (3), Delta; the adjoint sweep writes j_bar, alpha_bar, rho_bar over j, alpha, rho.
// the sample's geometry the frames of its splats need (geo_store / geo_load): written once per step-taking sample
// by the transfer, read by every element adjoint of the sample instead of rebuilding it in FP64 per (sample, splat)
constexpr int GEO_FIELDS = 16;
enum RecField { REC_J, REC_A = REC_J + 4, REC_RHO = REC_A + 4, REC_D = REC_RHO + 3, REC_GEO, REC_FIELDS = REC_GEO + GEO_FIELDS };
constexpr int BAR_FIELDS = 11;    // the adjoints of j (4), alpha (4), rho (3) of every step-taking sample

// A segment's composite transfer map: for its samples, front to back, P = O_k0 ... O_k1 and
// Sv = sum_k O_k0 ... O_{k-1} E_k (the RadiativeState of step.jl:342-398 started at the identity).
// In FP32 the Stokes sum is compensated (Kahan), so its error does not grow with the segment length.
template <class S>
struct SegMapT {
    S P[16], Sv[4], comp[4];
    KS_DEV void init() {
        for (int i = 0; i < 16; ++i) P[i] = S((i % 5 == 0) ? 1.0 : 0.0);
        for (int i = 0; i < 4; ++i) { Sv[i] = S(0.0); comp[i] = S(0.0); }
    }
    KS_DEV void advance(const S O[16], const S E[4]) {
        for (int i = 0; i < 4; ++i) add_stokes(P[4 * i] * E[0] + P[4 * i + 1] * E[1] + P[4 * i + 2] * E[2] + P[4 * i + 3] * E[3], i);
        S Pn[16];
        for (int i = 0; i < 4; ++i)
            for (int q = 0; q < 4; ++q) Pn[4 * i + q] = P[4 * i] * O[q] + P[4 * i + 1] * O[4 + q] + P[4 * i + 2] * O[8 + q] + P[4 * i + 3] * O[12 + q];
        for (int i = 0; i < 16; ++i) P[i] = Pn[i];
    }
    KS_DEV void add_stokes(const S pe, int i) {
        if constexpr (!std::is_same<S, f32>::value) {          // compensated in FP32 only
            Sv[i] += pe;
        } else {
            S y = pe - comp[i];
            S t = Sv[i] + y;
            comp[i] = (t - Sv[i]) - y;
            Sv[i] = t;
        }
    }
    // One sample's step without forming O: E = Dp sum_n e_n K'^n j, Sv += P E, then every row of P
    // becomes sum_n o_n (row K'^n) (products with K'^T), which keeps 4 + 16 values live instead of the
    // 64 of O, K'^2 and K'^3.
    KS_DEV void step(const S j4[4], const S a4[4], const S r3[3], S D) {
        S Dp, K6[6], o[4], e4[4];
        step_coefs(a4, r3, D, Dp, K6, o, e4);
        S u[4], y[4], E[4];
        for (int i = 0; i < 4; ++i) { u[i] = j4[i]; E[i] = e4[0] * u[i]; }
        for (int n = 1; n < 4; ++n) {
            kprime_mul(K6, u, y);
            for (int i = 0; i < 4; ++i) { u[i] = y[i]; E[i] += e4[n] * u[i]; }
        }
        for (int i = 0; i < 4; ++i) add_stokes(Dp * (P[4 * i] * E[0] + P[4 * i + 1] * E[1] + P[4 * i + 2] * E[2] + P[4 * i + 3] * E[3]), i);
        for (int i = 0; i < 4; ++i) {
            S v[4], acc[4];
            for (int q = 0; q < 4; ++q) { v[q] = P[4 * i + q]; acc[q] = o[0] * v[q]; }
            for (int n = 1; n < 4; ++n) {
                kprime_tmul(K6, v, y);
                for (int q = 0; q < 4; ++q) { v[q] = y[q]; acc[q] += o[n] * v[q]; }
            }
            for (int q = 0; q < 4; ++q) P[4 * i + q] = acc[q];
        }
    }
    KS_DEV void store(double* seg, int64_t m, int64_t nitems, int64_t item) const {
        double* o = seg + SEG_FIELDS * m * nitems + item;
        for (int q = 0; q < 16; ++q) o[q * nitems] = val(P[q]);
        for (int q = 0; q < 4; ++q) o[(16 + q) * nitems] = val(Sv[q]);
    }
};
using SegMap = SegMapT<double>;

// The half-orbit counter (WindingState, step.jl:362-398) over the valid samples outside the horizon
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
// ---- ks_k_geo.hip ----
hipError_t launch_root_case(hipStream_t s, double a, double thetao, const double* al, const double* be, int64_t npix, int8_t* numreals);
hipError_t launch_case_count(hipStream_t s, int nblocks, const int8_t* numreals, int64_t npix, int* bcount);
hipError_t launch_case_scatter(hipStream_t s, int nblocks, const int8_t* numreals, int64_t npix, const int* boff, int* perm);
hipError_t launch_pixel_constants(hipStream_t s, double a, double thetao, const double* al, const double* be, const int* perm, int64_t npix,
                                  double* soa);
hipError_t launch_march_store(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, int marcher, int nseg_max,
                              Samples smp, double* res, int fp32);
hipError_t launch_res_reduce(hipStream_t s, int64_t npix, int nseg_max, const double* res, double* rt, double* rp);
hipError_t launch_wind_cutoff(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, int nmax, double slab,
                              Samples smp, int* kstop);
hipError_t launch_samples_f32(hipStream_t s, int64_t n, const double* x, const double* y, const double* z, const double* t, float4* p4);
hipError_t launch_unsort4(hipStream_t s, const double* src, const int* perm, int64_t npix, int64_t nframes, double* dst);
hipError_t launch_unsort1(hipStream_t s, const double* src, const int* perm, int64_t npix, double* dst);

// ---- ks_k_transfer.hip ----
// How the transfer from stored samples is launched: samples per work item (chunk), one workgroup per task or a
// persistent grid over a work queue, and optionally a task list (the queue's order, or a sample of the work).
struct TransferLaunch {
    int chunk = CHUNK;
    int queue = 0;                 // 1: persistent workgroups taking tasks from a queue
    const int* tasks = nullptr;    // tasks to run, in this order (null: every task, in index order)
    int64_t ntasks = 0;            // the list's length
    int* head = nullptr;           // the queue head (one int of device memory; zeroed by the launch)
    uint32_t* cost = nullptr;      // queue form: per-task wall-clock ticks, indexed by task (null: not recorded)
    int group_scale = 1;           // queue form: persistent workgroups = resident workgroups x group_scale
};
int64_t transfer_tasks(int64_t npix, int nt, int nnu, int N, int chunk);
hipError_t launch_order_tasks(hipStream_t s, const uint32_t* cost, int64_t n, uint32_t* keys, int* ids, int* order, void* tmp, size_t& tmp_bytes);
hipError_t launch_fill_i32(hipStream_t s, int* p, int64_t n, int v);
hipError_t launch_chunk_transfer(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                 const SplatC* sp, int ns, const double* knots, const double* ktimes, int nknots,
                                 const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                 double* seg, double* rec, uint8_t* aflg, const TransferLaunch& tl);
hipError_t launch_chunk_transfer_f32(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                     const SplatCf* sp, int ns, const f32* knots, const f32* ktimes, int nknots,
                                     const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                     double* seg, f32* rec, uint8_t* aflg, const TransferLaunch& tl);
hipError_t launch_seg_polarized(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, int M,
                                const SplatC* sp, int ns, const double* knots, const double* ktimes, int nknots,
                                const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                double* seg, double* res);
hipError_t launch_ray_hits(hipStream_t s, double rp, int64_t npix, int N, Samples smp, const SplatC* sp, int ns, const double* times, int nt,
                           const int* kstop, uint8_t* hits, int* cnt);
hipError_t launch_ray_fill(hipStream_t s, int64_t items, int64_t npix, int ns, int cap, const uint8_t* hits, const int* iperm, int32_t* ids);
hipError_t launch_thin_parts(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, Samples smp,
                             const SplatC* sp, int ns, double t_obs, int nseg_max, double* part, double* res);
hipError_t launch_thin_combine(hipStream_t s, int64_t npix, int nseg_max, const double* part, double* out);
hipError_t launch_seg_combine(hipStream_t s, const double* soa, int64_t npix, int N, int M, int chunk, int nt, const double* nus, int nnu,
                              const double* seg, double* out);

// ---- ks_k_adjoint.hip ----
hipError_t launch_adj_combine(hipStream_t s, int64_t npix, int N, int chunk, int nt, const double* nus, int nnu, const double* seg,
                              const double* dstokes, double* wT, double* out);
hipError_t launch_adj_segment(hipStream_t s, int N, int chunk, int64_t nitems, const double* wT, const double* rec, double* bar,
                              const uint8_t* aflg, double* tails, Samples smp, int64_t npix);
hipError_t launch_adj_segment_f32(hipStream_t s, int N, int chunk, int64_t nitems, const double* wT, const f32* rec, f32* bar,
                                  const uint8_t* aflg, f32* tails, Samples smp, int64_t npix);
hipError_t launch_adj_splats_f32(hipStream_t s, int pops, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                 const SplatCf* sp, const SplatAdjf* sadj, int ns, const f32* knots, const f32* ktimes,
                                 const double* times, int nt, const double* nus, int64_t nitems, const f32* rec, const f32* bar,
                                 const uint8_t* aflg, int64_t unit0, int64_t nb, const double* ub, double* part, int* heads);
hipError_t launch_adj_splats(hipStream_t s, int pops, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                             const SplatC* sp, const SplatAdj* sadj, int ns, const double* knots, const double* ktimes,
                             const double* times, int nt, const double* nus, int64_t nitems, const double* rec, const double* bar,
                             const uint8_t* aflg, int64_t unit0, int64_t nb, const double* ub, double* part, int* heads);
hipError_t launch_adj_reduce(hipStream_t s, const double* part, int64_t nunits, int ns, int accumulate, double* dparams);
hipError_t launch_unit_bounds(hipStream_t s, int64_t npix, int N, Samples smp, const double* times, int nt, int64_t nitems,
                              const uint8_t* aflg, double* ub);

// ---- ks_k_vis.hip ----
hipError_t launch_vis_tables(hipStream_t s, int nuv, int na, int nb, double psize, const double* u, const double* v, double* A2, double* B2);
hipError_t launch_vis_gather(hipStream_t s, const double* out, const int* perm, int64_t npix, int na, int nb, int64_t f0, int nf, double* Im);
hipError_t launch_dgemm(hipStream_t s, bool transA, int M, int N, int K, const double* A, int64_t lda, int64_t sA, const double* B,
                        int64_t ldb, int64_t sB, double* C, int64_t ldc, int64_t sC, int batch);
hipError_t launch_dgemm_tn_splitk(hipStream_t s, int M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb,
                                  double* C, int nsplit, double* P);
constexpr int VIS_SPLITK = 8;      // K ranges of the seed product (n_alpha x 4 n_beta from 2 nuv)
hipError_t launch_vis_combine(hipStream_t s, int nuv, int nb, int nf, const double* C2, const double* B2, const double* tap, double scale, double* V);
hipError_t launch_vis_residual(hipStream_t s, int nuv, const double* V, const double* vis, const double* sigma, const double* gains,
                               const double* tap, double scale, double* wq, double* chi_k);
hipError_t launch_vis_rotate(hipStream_t s, int nuv, int nb, const double* wq, const double* B2, double* E2);
hipError_t launch_vis_scatter_seed(hipStream_t s, const double* Sd, const int* perm, int64_t npix, int na, int nb, int64_t f, int add, double* ads);
hipError_t launch_vis_scale_weights(hipStream_t s, int nuv, const double* w, const double* tap, double scale, double* wq);
hipError_t launch_image_residual(hipStream_t s, const double* out, const int* perm, int64_t npix, int64_t f, const double* data,
                                 const double* sigma, int sigma_per_pixel, const uint8_t* mask, double* ads, double* chi_p);
hipError_t launch_sum_fixed(hipStream_t s, const double* x, int64_t n, double* out);
hipError_t launch_closure_terms(hipStream_t s, int ntri, const int* tri, const double* cp_data, const double* cp_sig, int nquad, const int* quad,
                                const double* la_data, const double* la_sig, const double* V, double* chi_t, double* coef);
hipError_t launch_closure_weights(hipStream_t s, int nuv, int ntri, const int* off, const int* legs, const double* coef, const double* V, double* w);
hipError_t launch_bin_residual(hipStream_t s, const double* out, const int* iperm, int64_t npix, int64_t f, const int* off, const int* pts,
                               int64_t nbins, const double* data, const double* sigma, int sigma_per_pixel, const uint8_t* mask,
                               double* pseed, double* chi_p);
hipError_t launch_bin_seed(hipStream_t s, const int* perm, const int* pix, int64_t npix, int64_t f, const double* pseed, double* ads);

// ---- ks_k_tangent.hip ----
hipError_t launch_tangent(hipStream_t s, int kind, double a, double thetao, double lnL, const double* al, const double* be, int64_t npix, int N,
                          const SplatC* sp, int ns, const double* times, int nt, const double* nus, int nnu, int nmax, double slab,
                          double* out, double* dout);

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
// resident workgroups of `block` threads of a kernel on the current device (a persistent grid); 0 when unknown
template <class K>
inline int resident_groups(K kernel, int block) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess)
        return 0;
    return cus * (per > 0 ? per : 1);
}

}  // namespace ksd
