// libkerrsplat_hip: the C-ABI declared in include/kerrsplat_hip.h over the MI355X (gfx950) kernels of
// KerrSplat's geodesic march + polarized transfer (declared in ks_kernels.hpp).
//
// A cache holds one camera's case-sorted per-pixel constants (structure of arrays), optionally its
// stored samples, and reusable device buffers; a model holds one splat population's per-splat
// constants. Every call that returns host data synchronizes the context's stream.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>
#include <atomic>
#include <mutex>

#include "ks_kernels.hpp"
#include "ks_hostsplat.hpp"
#include "../../include/kerrsplat_hip.h"

using namespace ksd;

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
uint64_t next_model_id() {
    static std::atomic<uint64_t> n{0};
    return ++n;
}
// a grow-only device buffer
template <class T> struct Buf {
    T* p = nullptr;
    size_t cap = 0;
    int grow(size_t n) {
        if (cap >= n && p) return 0;
        dfree(p); cap = 0;
        if (dalloc(&p, n)) return KS_ERR_HIP;
        cap = n;
        return 0;
    }
    void release() { dfree(p); cap = 0; }
};

}  // namespace

// The workspace of ks_visibilities_noalloc: one fixed array per GPU, held until the process ends. In host code
// `scratch` names a host-side stand-in, not device memory: its device address on each GPU is taken once, with
// that GPU current, by hipGetSymbolAddress (scratch_address) and only that address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[KS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int KS_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[KS_MAX_DEVICES] = {};

// the device address of `scratch` on `device` (which must be current), recorded on first use
int scratch_address(int device, void** out) {
    if (device < 0 || device >= KS_MAX_DEVICES) return fail(KS_ERR_ARG, "device %d beyond the %d the library tracks", device, KS_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(KS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return KS_OK;
}
}  // namespace

struct ks_ctx {
    int device = 0;
    void* scratch_dev = nullptr;                       // device address of `scratch` on this device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int NSTAGE = 5;
    hipEvent_t evs[NSTAGE][2] = {};                // around the last launch of each stage
    bool staged[NSTAGE] = {};
    bool capturing = false;                        // a fit graph is being captured on the stream: no events, no host reads
    hipError_t stage_begin(int k) { return capturing ? hipSuccess : hipEventRecord(evs[k][0], stream); }
    hipError_t stage_end(int k) {
        if (capturing) return hipSuccess;
        staged[k] = true;
        return hipEventRecord(evs[k][1], stream);
    }
};

struct ks_cache {
    ks_ctx* ctx = nullptr;
    int64_t npix = 0;
    int N = 0;
    int store = 0;
    int nalpha = 0, nbeta = 1;
    double spin = NAN, thetao = NAN;
    int marcher = KS_MARCHER_DIRECT, M = 64;
    int precision = KS_PRECISION_FP64;                 // of the transfer and the adjoint over stored samples
    int march_precision = KS_PRECISION_FP64;           // of the Recurrence march into stored samples (mixed: FP32 bulk)
    int generation = 0;
    int64_t counts[3] = {0, 0, 0};
    double *al = nullptr, *be = nullptr;
    int8_t* numreals = nullptr;
    int* perm = nullptr;
    int* tmap = nullptr;                               // wavefront grouping of the sorted pixels (patches of each screen part)
    int* iperm = nullptr;                              // screen index -> sorted slot
    int* bcount = nullptr; int nblocks = 0;
    double* soa = nullptr;
    double *st = nullptr, *sr = nullptr, *sth = nullptr, *sph = nullptr; uint8_t* sfl = nullptr;
    double *sx = nullptr, *sy = nullptr, *sz = nullptr;        // quasi-Cartesian position of each stored sample
    float4* sp4 = nullptr;                                     // their FP32 copy with t~ (x, y, z, t) for the FP32 element adjoints
    int64_t march_version = 0, p4_version = -1;                // marches into the samples; the march the copy was made from
    double* thin_part = nullptr;                               // per-segment thin sums (max_segments x npix)
    double *rt = nullptr, *rp = nullptr;
    double* out = nullptr; int64_t out_frames = 0;    // sorted-order Stokes images
    double* out_screen = nullptr;
    double* img1 = nullptr;                            // thin image (sorted)
    // visibility and χ² workspace (grow-only): (u, v, taper, gains, data, sigma), the phase tables
    // A2 (2 nuv x n_alpha) and B2 (2 nuv x n_beta), the images as n_alpha x 4 n_beta matrices, the
    // partial products C2 (2 nuv x 4 n_beta per frame), V, the rotated weights E2, the screen seed,
    // per-baseline or per-pixel χ² terms and their sum, the image data, noise and mask
    Buf<double> vu, vv, vt, vg, vd, vs, vA2, vB2, vIm, vC2, vV, vwq, vE2, vSd, vSp, vchi, vsum, idata, isig;
    Buf<uint8_t> imask;
    Buf<double> bseed;                                 // binned χ²: the seed of every binned pixel
    Buf<int> ctri, cquad, coff, clegs;                 // closures: triangles, quadrangles, per-baseline term lists
    Buf<double> cdat, cterm;                           // closures: data and sigmas; per-term χ² and coefficients
    double* seg = nullptr; int64_t seg_cap = 0;        // per-segment transfer maps
    double* dtn = nullptr; int64_t dtn_cap = 0;        // observation times then frequencies (device)
    std::vector<double> htn;                           // their host staging copy (outlives the async copy)
    double* segres = nullptr;                          // per-segment anchor residuals (2 x nseg_max x npix)
    int* kstop = nullptr;                              // half-orbit cutoff per pixel
    // adjoint workspace (grow-only): records, flags, tails, per-segment (w_in, T_out), seeds, partial sums
    double* arec = nullptr; double* abar = nullptr; uint8_t* aflg = nullptr; double* atail = nullptr; double* awT = nullptr; double* ads = nullptr;
    double* apart = nullptr; double* adpar = nullptr; double* aub = nullptr;
    int64_t a_items = 0, a_part = 0, a_dpar = 0, a_ub = 0, a_wT = 0;
    // the last forward pass of the dual sweep (ks_gradient_forward): what ks_gradient_reverse needs
    uint64_t fw_model = 0;                             // id of its model (0: no forward pass to reverse)
    int fw_nt = 0, fw_nnu = 0, fw_generation = -1, fw_precision = -1;
    std::vector<uint8_t> fw_seeded;                    // frames whose device seed a *_chi2_seed call wrote
    int fw_chunk = CHUNK;                              // samples per chunk of its segment maps
    std::vector<int> perm_host;
    // launch choices of the transfer over stored samples (ks_cache_tuning, ks_autotune): samples per work item,
    // the persistent work-queue form, the wavefront patch (patch_a x patch_b points of a screen part)
    int chunk = CHUNK, queue = 0, patch_a = 8, patch_b = 8;
    // screen parts of the point list (ks_cache_set_parts): consecutive runs of points, each an n_alpha x n_beta
    // screen in screen order; empty: the whole camera is one nalpha x nbeta screen
    std::vector<int64_t> part_off; std::vector<int> part_na, part_nb;
    // the work queue: its head, the per-task ticks of the last queue launch, the cost-descending order made from
    // them and the launch shape that order belongs to, the sort's work arrays
    Buf<int> qhead, qorder, qids, aheads;              // aheads: the element adjoints' per-XCD queue heads
    Buf<uint32_t> qcost, qkeys;
    Buf<uint8_t> qtmp;
    int64_t q_shape[6] = {-1, -1, -1, -1, -1, -1};
    bool q_ordered = false;
};

// Binning (camera.jl:62-80) on the device: the points of every pixel as a list in ascending screen order
// (off[q] .. off[q+1]), and the pixel of every point
struct ks_binning {
    ks_ctx* ctx = nullptr;
    int64_t npoints = 0, nbins = 0;
    int na = 0, nb = 0;
    int* off = nullptr; int* pts = nullptr; int* pix = nullptr;
};

struct ks_model {
    ks_ctx* ctx = nullptr;
    uint64_t id = 0;                                   // unique per created model: a model freed and re-created at the same
                                                       // address is not the model of an earlier forward pass
    int kind = KS_MODEL_THERMAL, nrows = 0, ns = 0;
    std::vector<int> hrows;                            // parameter rows of every splat (21 thermal/knot, 22 power law/kappa)
    int pops = 0;                                      // the populations it holds (POP_THERMAL, POP_POWERLAW, POP_KAPPA, POP_KNOTS)
    SplatC* d_splats = nullptr;
    SplatAdj* d_adj = nullptr;                         // thermal splats: constants of the adjoint
    double* d_knots = nullptr; double* d_ktimes = nullptr; int nknots = 0;
    int64_t nkn = 0, nkt = 0;                          // knot triples and knot times held (knot models and composites with knot parts)
    // the same constants in FP32 (the mixed-precision transfer and adjoint)
    SplatCf* d_splats_f = nullptr; SplatAdjf* d_adj_f = nullptr; f32* d_knots_f = nullptr; f32* d_ktimes_f = nullptr;
    double gamma_max = 0.0;                            // of a single-population model: what its constants were made with
    std::vector<double> h_knots;                       // and its knots (the per-splat constants bound the knot paths)
};

extern "C" {

const char* ks_last_error(void) { return g_err.c_str(); }

int ks_ctx_create(int device, ks_ctx** out) {
    if (!out) return fail(KS_ERR_ARG, "ks_ctx_create: null output pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(KS_ERR_ARG, "device %d outside 0:%d", device, n - 1);
    HIPCHK(hipSetDevice(device));
    void* scr = nullptr;
    if (int rc = scratch_address(device, &scr)) return rc;
    ks_ctx* c = new ks_ctx;
    c->device = device;
    c->scratch_dev = scr;
    HIPCHK(hipStreamCreate(&c->stream));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    for (int k = 0; k < ks_ctx::NSTAGE; ++k) { HIPCHK(hipEventCreate(&c->evs[k][0])); HIPCHK(hipEventCreate(&c->evs[k][1])); }
    *out = c;
    return KS_OK;
}

int ks_ctx_destroy(ks_ctx* c) {
    if (!c) return KS_OK;
    (void)hipSetDevice(c->device);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int k = 0; k < ks_ctx::NSTAGE; ++k)
        for (int q = 0; q < 2; ++q) if (c->evs[k][q]) (void)hipEventDestroy(c->evs[k][q]);
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
int ks_stage_ms(ks_ctx* c, int stage, float* ms) {
    if (!c || !ms) return fail(KS_ERR_ARG, "null argument");
    if (stage < 0 || stage >= ks_ctx::NSTAGE) return fail(KS_ERR_ARG, "unknown stage %d", stage);
    if (!c->staged[stage]) return fail(KS_ERR_STATE, "stage %d has not run on this context yet", stage);
    HIPCHK(hipEventSynchronize(c->evs[stage][1]));
    HIPCHK(hipEventElapsedTime(ms, c->evs[stage][0], c->evs[stage][1]));
    return KS_OK;
}

int ks_cache_create(ks_ctx* ctx, int64_t npix, int N, int store_samples, ks_cache** out) {
    if (!ctx || !out) return fail(KS_ERR_ARG, "ks_cache_create: null argument");
    if (N < 1) return fail(KS_ERR_ARG, "need at least one sample per ray");
    if (N > (1 << 24)) return fail(KS_ERR_ARG, "N = %d above the supported 2^24 samples per ray", N);
    if (npix < 1) return fail(KS_ERR_ARG, "need at least one pixel");
    if (npix > ((int64_t)1 << 30)) return fail(KS_ERR_ARG, "npix %lld above the supported 2^30", (long long)npix);
    HIPCHK(hipSetDevice(ctx->device));
    ks_cache* c = new ks_cache;
    c->ctx = ctx; c->npix = npix; c->N = N; c->store = store_samples ? 1 : 0;
    c->nalpha = (int)npix; c->nbeta = 1;
    c->nblocks = (int)grid_for(npix, PART_BLOCK);
    int rc = 0;
    rc |= dalloc(&c->al, npix); rc |= dalloc(&c->be, npix); rc |= dalloc(&c->numreals, npix); rc |= dalloc(&c->perm, npix);
    rc |= dalloc(&c->tmap, npix);
    rc |= dalloc(&c->iperm, npix);
    rc |= dalloc(&c->bcount, 3 * (size_t)c->nblocks);
    rc |= dalloc(&c->soa, (size_t)PF_COUNT * npix);
    rc |= dalloc(&c->rt, npix); rc |= dalloc(&c->rp, npix);
    rc |= dalloc(&c->segres, (size_t)2 * max_segments(N) * npix);
    if (c->store) {
        size_t ns = (size_t)npix * N;
        rc |= dalloc(&c->st, ns); rc |= dalloc(&c->sr, ns); rc |= dalloc(&c->sth, ns); rc |= dalloc(&c->sph, ns); rc |= dalloc(&c->sfl, ns);
        rc |= dalloc(&c->sx, ns); rc |= dalloc(&c->sy, ns); rc |= dalloc(&c->sz, ns);
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
    dfree(c->al); dfree(c->be); dfree(c->numreals); dfree(c->perm); dfree(c->tmap); dfree(c->iperm); dfree(c->bcount); dfree(c->soa);
    dfree(c->st); dfree(c->sr); dfree(c->sth); dfree(c->sph); dfree(c->sfl); dfree(c->rt); dfree(c->rp);
    dfree(c->sx); dfree(c->sy); dfree(c->sz); dfree(c->thin_part); dfree(c->sp4);
    dfree(c->out); dfree(c->out_screen); dfree(c->img1);
    for (Buf<double>* b : {&c->vu, &c->vv, &c->vt, &c->vg, &c->vd, &c->vs, &c->vA2, &c->vB2, &c->vIm, &c->vC2, &c->vV, &c->vwq, &c->vE2,
                           &c->vSd, &c->vSp, &c->vchi, &c->vsum, &c->idata, &c->isig, &c->bseed})
        b->release();
    c->imask.release();
    c->ctri.release(); c->cquad.release(); c->coff.release(); c->clegs.release(); c->cdat.release(); c->cterm.release();
    dfree(c->seg); dfree(c->segres); dfree(c->kstop); dfree(c->dtn);
    dfree(c->arec); dfree(c->abar); dfree(c->aflg); dfree(c->atail); dfree(c->awT); dfree(c->ads); dfree(c->apart); dfree(c->adpar); dfree(c->aub);
    c->qhead.release(); c->qorder.release(); c->qids.release(); c->aheads.release(); c->qcost.release(); c->qkeys.release(); c->qtmp.release();
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

static Samples samples_of(const ks_cache* c) {
    Samples S;
    S.t = c->st; S.r = c->sr; S.th = c->sth; S.ph = c->sph; S.x = c->sx; S.y = c->sy; S.z = c->sz; S.fl = c->sfl;
    S.tmap = c->tmap;
    S.p4 = c->p4_version == c->march_version ? c->sp4 : nullptr;
    return S;
}
// the FP32 copy of the stored positions and times (for the FP32 element adjoints), made once per march
static int ensure_p4(ks_cache* c) {
    if (c->sp4 && c->p4_version == c->march_version) return KS_OK;
    const int64_t n = c->npix * (int64_t)c->N;
    if (!c->sp4 && dalloc(&c->sp4, (size_t)n)) return KS_ERR_HIP;
    HIPCHK(launch_samples_f32(c->ctx->stream, n, c->sx, c->sy, c->sz, c->st, c->sp4));
    c->p4_version = c->march_version;
    return KS_OK;
}
static Samples no_samples() {
    Samples S;
    S.t = S.r = S.th = S.ph = S.x = S.y = S.z = nullptr; S.fl = nullptr;
    S.tmap = nullptr;
    return S;
}
// the cache's marcher into its sample storage (march!, cache.jl:182-190), anchor segments in parallel
static int launch_store_march(ks_cache* c) {
    ks_ctx* x = c->ctx;
    const int nseg = max_segments(c->N);
    const int M = c->marcher == KS_MARCHER_DIRECT ? 64 : c->M;
    HIPCHK(x->stage_begin(KS_STAGE_MARCH));
    HIPCHK(launch_march_store(x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, M, c->marcher == KS_MARCHER_DIRECT ? 0 : 1, nseg,
                              samples_of(c), c->segres, c->march_precision == KS_PRECISION_MIXED ? 1 : 0));
    HIPCHK(x->stage_end(KS_STAGE_MARCH));
    c->march_version += 1;
    HIPCHK(launch_res_reduce(x->stream, c->npix, nseg, c->segres, c->rt, c->rp));
    return KS_OK;
}
static bool has_samples(const ks_cache* c) { return c->store && c->marcher != KS_MARCHER_FUSED && c->generation >= 1; }

// screen index -> sorted slot
static std::vector<int> iperm_of(const ks_cache* c) {
    std::vector<int> inv(c->npix);
    for (int64_t j = 0; j < c->npix; ++j) inv[c->perm_host[j]] = (int)j;
    return inv;
}
// Wavefront grouping for the kernels over stored samples: the sorted pixels in patch_a x patch_b patches of each
// screen part, so that the 64 lanes of a wavefront look at one compact patch of sky (a splat then covers most or none
// of them); a patch row stays contiguous in the case-sorted order.
static int build_tmap(ks_cache* c) {
    const std::vector<int> inv = iperm_of(c);
    std::vector<int64_t> off = c->part_off;
    std::vector<int> na = c->part_na, nb = c->part_nb;
    if (off.empty()) { off = {0}; na = {c->nalpha}; nb = {c->nbeta}; }
    std::vector<int> tm;
    tm.reserve(c->npix);
    const int pa = c->patch_a, pb = c->patch_b;
    for (size_t q = 0; q < off.size(); ++q)
        for (int tb = 0; tb < nb[q]; tb += pb)
            for (int ta = 0; ta < na[q]; ta += pa)
                for (int ib = tb; ib < std::min(tb + pb, nb[q]); ++ib)
                    for (int ia = ta; ia < std::min(ta + pa, na[q]); ++ia) tm.push_back(inv[off[q] + ia + (int64_t)na[q] * ib]);
    if ((int64_t)tm.size() != c->npix) return fail(KS_ERR_STATE, "screen parts cover %zu of %lld points", tm.size(), (long long)c->npix);
    HIPCHK(hipMemcpy(c->tmap, tm.data(), c->npix * sizeof(int), hipMemcpyHostToDevice));
    c->q_ordered = false;          // the tasks hold other items now
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
    // negative spins are the y -> -y mirror of positive ones (the reference's reflection gate runs at a = -0.94);
    // a = 0 is excluded: the closed-form roots and the polar integrals divide by a
    if (!(fabs(a) > 0.0 && fabs(a) < 1.0)) return fail(KS_ERR_ARG, "spin a = %g outside 0 < |a| < 1", a);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    hipStream_t s = x->stream;
    const int64_t n = c->npix;
    c->nalpha = nalpha; c->nbeta = nbeta;
    HIPCHK(hipMemcpyAsync(c->al, alpha, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->be, beta, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(launch_root_case(s, a, thetao, c->al, c->be, n, c->numreals));
    HIPCHK(launch_case_count(s, c->nblocks, c->numreals, n, c->bcount));
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
    HIPCHK(launch_case_scatter(s, c->nblocks, c->numreals, n, c->bcount, c->perm));
    HIPCHK(launch_pixel_constants(s, a, thetao, c->al, c->be, c->perm, n, c->soa));
    c->spin = a; c->thetao = thetao; c->marcher = marcher; c->M = M;
    if (marcher != KS_MARCHER_FUSED) {
        int rc = launch_store_march(c);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->perm_host.data(), c->perm, n * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(c->iperm, iperm_of(c).data(), c->npix * sizeof(int), hipMemcpyHostToDevice));
    if (int rc = build_tmap(c)) return rc;
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
int ks_model_create(ks_ctx* ctx, int kind, int nrows, int ns, const double* params, double gamma_max,
                    const double* knots, const double* knot_times, int nknots, ks_model** out) {
    if (!ctx || !out) return fail(KS_ERR_ARG, "ks_model_create: null argument");
    int want = kind == KS_MODEL_THIN ? 14 : (kind == KS_MODEL_THERMAL || kind == KS_MODEL_KNOT) ? 21 : (kind == KS_MODEL_POWERLAW || kind == KS_MODEL_KAPPA) ? 22 : -1;
    if (want < 0) return fail(KS_ERR_ARG, "unknown model kind %d", kind);
    if (nrows != want) return fail(KS_ERR_DIM, "model kind %d takes %d parameter rows, got %d", kind, want, nrows);
    if (ns < 0 || (ns > 0 && !params)) return fail(KS_ERR_ARG, "bad splat count or parameters");
    if (kind == KS_MODEL_KNOT && (nknots < 1 || !knots || !knot_times)) return fail(KS_ERR_ARG, "knot splats need knots and knot times");
    // centres that follow knots: knot splats, or thermal, power-law and kappa splats given knots
    const bool with_knots = kind != KS_MODEL_THIN && knots && knot_times && nknots >= 1;
    if (with_knots)
        for (int k = 0; k + 1 < nknots; ++k)
            if (!(knot_times[k + 1] > knot_times[k])) return fail(KS_ERR_ARG, "knot times must increase");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<SplatC> h(ns > 0 ? ns : 1);
    std::vector<SplatAdj> ha(ns > 0 ? ns : 1);
    for (int i = 0; i < ns; ++i)
        host_splat_constants(kind, nrows, i, params + (size_t)nrows * i, gamma_max, with_knots, knots, nknots, h[i], ha[i]);
    ks_model* m = new ks_model;
    m->id = next_model_id();
    m->ctx = ctx; m->kind = kind; m->nrows = nrows; m->ns = ns;
    m->hrows.assign((size_t)ns, nrows);
    m->pops = (kind == KS_MODEL_POWERLAW ? POP_POWERLAW : kind == KS_MODEL_KAPPA ? POP_KAPPA : POP_THERMAL) | (with_knots ? POP_KNOTS : 0);
    m->gamma_max = gamma_max;
    if (with_knots) m->h_knots.assign(knots, knots + (size_t)3 * nknots * ns);
    if (dalloc(&m->d_splats, h.size()) || dalloc(&m->d_adj, ha.size()) || dalloc(&m->d_splats_f, h.size()) || dalloc(&m->d_adj_f, ha.size())) {
        ks_model_destroy(m); return KS_ERR_HIP;
    }
    HIPCHK(hipMemcpy(m->d_splats, h.data(), h.size() * sizeof(SplatC), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->d_adj, ha.data(), ha.size() * sizeof(SplatAdj), hipMemcpyHostToDevice));
    {
        std::vector<SplatCf> hf(h.size());
        std::vector<SplatAdjf> haf(ha.size());
        for (size_t i = 0; i < h.size(); ++i) { splat_as(h[i], hf[i]); adj_as(ha[i], haf[i]); }
        HIPCHK(hipMemcpy(m->d_splats_f, hf.data(), hf.size() * sizeof(SplatCf), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_adj_f, haf.data(), haf.size() * sizeof(SplatAdjf), hipMemcpyHostToDevice));
    }
    if (with_knots) {
        m->nknots = nknots;
        m->nkn = (int64_t)nknots * ns; m->nkt = nknots;
        if (dalloc(&m->d_knots, (size_t)3 * nknots * ns) || dalloc(&m->d_ktimes, (size_t)nknots)) { ks_model_destroy(m); return KS_ERR_HIP; }
        HIPCHK(hipMemcpy(m->d_knots, knots, (size_t)3 * nknots * ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_ktimes, knot_times, (size_t)nknots * sizeof(double), hipMemcpyHostToDevice));
        std::vector<f32> kf((size_t)3 * nknots * ns), tf((size_t)nknots);
        for (size_t i = 0; i < kf.size(); ++i) kf[i] = f32(knots[i]);
        for (size_t i = 0; i < tf.size(); ++i) tf[i] = f32(knot_times[i]);
        if (dalloc(&m->d_knots_f, kf.size()) || dalloc(&m->d_ktimes_f, tf.size())) { ks_model_destroy(m); return KS_ERR_HIP; }
        HIPCHK(hipMemcpy(m->d_knots_f, kf.data(), kf.size() * sizeof(f32), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->d_ktimes_f, tf.data(), tf.size() * sizeof(f32), hipMemcpyHostToDevice));
    }
    *out = m;
    return KS_OK;
}

static_assert(KS_MODEL_THIN == MODEL_THIN && KS_MODEL_THERMAL == MODEL_THERMAL && KS_MODEL_POWERLAW == MODEL_POWERLAW &&
              KS_MODEL_KAPPA == MODEL_KAPPA && KS_MODEL_KNOT == MODEL_KNOT && KS_MODEL_COMPOSITE == MODEL_COMPOSITE,
              "the C-ABI model kinds are the device population tags");

int ks_model_composite(ks_ctx* ctx, int nparts, const ks_model* const* parts, ks_model** out) {
    if (!ctx || !out || !parts || nparts < 1) return fail(KS_ERR_ARG, "ks_model_composite: null argument or no parts");
    int64_t ns = 0, nkn = 0, nkt = 0;    // splats, knot triples and knot times of all parts
    for (int p = 0; p < nparts; ++p) {
        const ks_model* q = parts[p];
        if (!q) return fail(KS_ERR_ARG, "ks_model_composite: part %d is null", p);
        if (q->ctx != ctx) return fail(KS_ERR_ARG, "ks_model_composite: part %d belongs to another context", p);
        if (q->kind != KS_MODEL_THERMAL && q->kind != KS_MODEL_POWERLAW && q->kind != KS_MODEL_KAPPA && q->kind != KS_MODEL_KNOT &&
            q->kind != KS_MODEL_COMPOSITE)
            return fail(KS_ERR_ARG, "ks_model_composite: part %d has kind %d (parts are thermal, power-law, kappa, knot or composite)", p,
                        q->kind);
        ns += q->ns;
        nkn += q->nkn; nkt += q->nkt;
    }
    if (ns > INT32_MAX || nkn > INT32_MAX) return fail(KS_ERR_DIM, "ks_model_composite: %lld splats, %lld knots", (long long)ns, (long long)nkn);
    HIPCHK(hipSetDevice(ctx->device));
    ks_model* m = new ks_model;
    m->id = next_model_id();
    m->ctx = ctx; m->kind = KS_MODEL_COMPOSITE; m->nrows = 0; m->ns = (int)ns;
    for (int p = 0; p < nparts; ++p) {
        m->hrows.insert(m->hrows.end(), parts[p]->hrows.begin(), parts[p]->hrows.end());
        m->pops |= parts[p]->pops;
    }
    const size_t n = ns > 0 ? (size_t)ns : 1;
    if (dalloc(&m->d_splats, n) || dalloc(&m->d_adj, n) || dalloc(&m->d_splats_f, n) || dalloc(&m->d_adj_f, n)) {
        ks_model_destroy(m); return KS_ERR_HIP;
    }
#define HIPCHK_M(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { ks_model_destroy(m); return fail(KS_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
    // the splats (with each knot splat's knot0 / kt0 moved by the offsets of its part's knots) and the knots, concatenated
    std::vector<SplatC> h(n);
    std::vector<SplatCf> hf(n);
    std::vector<double> kn((size_t)3 * nkn), kt((size_t)nkt);
    size_t off = 0, okn = 0, okt = 0;
    for (int p = 0; p < nparts; ++p) {
        const ks_model* q = parts[p];
        if (q->ns > 0) {
            HIPCHK_M(hipMemcpy(h.data() + off, q->d_splats, q->ns * sizeof(SplatC), hipMemcpyDeviceToHost));
            HIPCHK_M(hipMemcpy(m->d_adj + off, q->d_adj, q->ns * sizeof(SplatAdj), hipMemcpyDeviceToDevice));
            HIPCHK_M(hipMemcpy(m->d_adj_f + off, q->d_adj_f, q->ns * sizeof(SplatAdjf), hipMemcpyDeviceToDevice));
            for (int i = 0; i < q->ns; ++i)
                if (h[off + i].nk > 0) { h[off + i].knot0 += (int)okn; h[off + i].kt0 += (int)okt; }
        }
        if (q->nkn > 0) {
            HIPCHK_M(hipMemcpy(kn.data() + 3 * okn, q->d_knots, (size_t)3 * q->nkn * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK_M(hipMemcpy(kt.data() + okt, q->d_ktimes, (size_t)q->nkt * sizeof(double), hipMemcpyDeviceToHost));
        }
        off += q->ns; okn += q->nkn; okt += q->nkt;
    }
    for (size_t i = 0; i < n; ++i) splat_as(h[i], hf[i]);
    HIPCHK_M(hipMemcpy(m->d_splats, h.data(), n * sizeof(SplatC), hipMemcpyHostToDevice));
    HIPCHK_M(hipMemcpy(m->d_splats_f, hf.data(), n * sizeof(SplatCf), hipMemcpyHostToDevice));
    if (nkn > 0) {
        m->nknots = (int)nkt; m->nkn = nkn; m->nkt = nkt;
        std::vector<f32> knf(kn.size()), ktf(kt.size());
        for (size_t i = 0; i < kn.size(); ++i) knf[i] = f32(kn[i]);
        for (size_t i = 0; i < kt.size(); ++i) ktf[i] = f32(kt[i]);
        if (dalloc(&m->d_knots, kn.size()) || dalloc(&m->d_ktimes, kt.size()) || dalloc(&m->d_knots_f, kn.size()) || dalloc(&m->d_ktimes_f, kt.size())) {
            ks_model_destroy(m); return KS_ERR_HIP;
        }
        HIPCHK_M(hipMemcpy(m->d_knots, kn.data(), kn.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK_M(hipMemcpy(m->d_ktimes, kt.data(), kt.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK_M(hipMemcpy(m->d_knots_f, knf.data(), knf.size() * sizeof(f32), hipMemcpyHostToDevice));
        HIPCHK_M(hipMemcpy(m->d_ktimes_f, ktf.data(), ktf.size() * sizeof(f32), hipMemcpyHostToDevice));
    }
    *out = m;
    return KS_OK;
#undef HIPCHK_M
}

int ks_model_destroy(ks_model* m) {
    if (!m) return KS_OK;
    (void)hipSetDevice(m->ctx->device);
    dfree(m->d_splats); dfree(m->d_adj); dfree(m->d_knots); dfree(m->d_ktimes);
    dfree(m->d_splats_f); dfree(m->d_adj_f); dfree(m->d_knots_f); dfree(m->d_ktimes_f);
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

// Segment maps per item of the transfer: chunks of the stored samples or the pixel's anchor segments
static int seg_count(const ks_cache* c, int chunk) { return std::max(max_segments(c->N), (c->N + chunk - 1) / chunk); }

// The launch of the transfer over stored samples as the cache's tuning asks: one workgroup per task, or the
// persistent work queue. The queue takes the tasks in the cost-descending order of an earlier launch of the same
// shape when there is one, records every task's ticks, and orders the tasks by them for the next launch.
static int transfer_launch(ks_cache* c, const ks_model* m, const double* d_tn, int nt, int nnu, double L, const int* kst, void* rec,
                           uint8_t* aflg, const TransferLaunch* sample) {
    ks_ctx* x = c->ctx;
    TransferLaunch tl;
    tl.chunk = c->chunk;
    tl.queue = (c->queue & KS_QUEUE_TRANSFER) ? 1 : 0;
    const int64_t ntasks = transfer_tasks(c->npix, nt, nnu, c->N, c->chunk);
    bool order_after = false;
    if (sample) {
        tl.tasks = sample->tasks; tl.ntasks = sample->ntasks;
    }
    if (tl.queue) {
        if (c->qhead.grow(1)) return KS_ERR_HIP;
        tl.head = c->qhead.p;
        if (!sample && x->capturing) {
            // a captured launch takes the order the last eager launch of this shape left, and records nothing
            const int64_t shape[6] = {ntasks, c->chunk, (int64_t)m->id, c->precision, rec ? 1 : 0, (int64_t)nt * 65536 + nnu};
            if (c->q_ordered && std::equal(shape, shape + 6, c->q_shape)) { tl.tasks = c->qorder.p; tl.ntasks = ntasks; }
        } else if (!sample) {
            const int64_t shape[6] = {ntasks, c->chunk, (int64_t)m->id, c->precision, rec ? 1 : 0, (int64_t)nt * 65536 + nnu};
            if (!std::equal(shape, shape + 6, c->q_shape)) { std::copy(shape, shape + 6, c->q_shape); c->q_ordered = false; }
            if (c->qcost.grow((size_t)ntasks)) return KS_ERR_HIP;
            tl.cost = c->qcost.p;
            if (c->q_ordered) { tl.tasks = c->qorder.p; tl.ntasks = ntasks; }
            order_after = true;
        }
    }
    const int kind = m->kind;
    if (c->precision == KS_PRECISION_MIXED)
        HIPCHK(launch_chunk_transfer_f32(x->stream, kind, c->spin, c->thetao, c->soa, c->npix, c->N, samples_of(c), m->d_splats_f, m->ns,
                                         m->d_knots_f, m->d_ktimes_f, m->nknots, d_tn, nt, d_tn + nt, nnu, L, kst, c->seg, (f32*)rec, aflg, tl));
    else
        HIPCHK(launch_chunk_transfer(x->stream, kind, c->spin, c->thetao, c->soa, c->npix, c->N, samples_of(c), m->d_splats, m->ns,
                                     m->d_knots, m->d_ktimes, m->nknots, d_tn, nt, d_tn + nt, nnu, L, kst, c->seg, (double*)rec, aflg, tl));
    if (order_after) {
        size_t bytes = 0;
        HIPCHK(launch_order_tasks(x->stream, c->qcost.p, ntasks, nullptr, nullptr, nullptr, nullptr, bytes));
        if (c->qorder.grow((size_t)ntasks) || c->qids.grow((size_t)ntasks) || c->qkeys.grow((size_t)ntasks) || c->qtmp.grow(std::max<size_t>(bytes, 1)))
            return KS_ERR_HIP;
        // the order for the next launch (stream order: the sort starts after the launch that reads qorder has ended)
        HIPCHK(launch_order_tasks(x->stream, c->qcost.p, ntasks, c->qkeys.p, c->qids.p, c->qorder.p, c->qtmp.p, bytes));
        c->q_ordered = true;
    }
    return KS_OK;
}

// The transfer of every (pixel, time, frequency) item into the cache's segment maps: from the stored
// samples when the cache holds them (one work item per chunk of samples), else fused with the march (one
// work item per anchor segment). Records for the adjoint when rec != nullptr (stored samples only).
static int launch_transfer(ks_cache* c, const ks_model* m, const double* d_tn, int nt, int nnu, double L, int nmax, double slab,
                           double* rec, uint8_t* aflg) {
    ks_ctx* x = c->ctx;
    const int64_t items = c->npix * nt * nnu;
    const int nseg = seg_count(c, c->chunk);
    if (c->seg_cap < (int64_t)SEG_FIELDS * nseg * items) {
        dfree(c->seg);
        if (dalloc(&c->seg, (size_t)SEG_FIELDS * nseg * items)) return KS_ERR_HIP;
        c->seg_cap = (int64_t)SEG_FIELDS * nseg * items;
    }
    const int* kst = nullptr;
    if (nmax >= 0) {
        if (!c->kstop) { if (dalloc(&c->kstop, (size_t)c->npix)) return KS_ERR_HIP; }
        HIPCHK(launch_wind_cutoff(x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, c->M, nmax, slab,
                                  has_samples(c) ? samples_of(c) : no_samples(), c->kstop));
        kst = c->kstop;
    }
    HIPCHK(x->stage_begin(KS_STAGE_TRANSFER));
    if (has_samples(c)) {
        if (int rc = transfer_launch(c, m, d_tn, nt, nnu, L, kst, rec, aflg, nullptr)) return rc;
    } else
        HIPCHK(launch_seg_polarized(x->stream, m->kind, c->spin, c->thetao, c->soa, c->npix, c->N, c->M, m->d_splats, m->ns,
                                    m->d_knots, m->d_ktimes, m->nknots, d_tn, nt, d_tn + nt, nnu, L, kst, c->seg, c->segres));
    HIPCHK(x->stage_end(KS_STAGE_TRANSFER));
    if (!has_samples(c)) {
        HIPCHK(launch_res_reduce(x->stream, c->npix, max_segments(c->N), c->segres, c->rt, c->rp));
    }
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
    c->fw_model = 0;                   // the render reuses the segment maps, times and images of a forward pass
    const int64_t frames = (int64_t)nt * nnu;
    if (ensure_out(c, frames)) return KS_ERR_HIP;
    const double* d_tn = stage_times(c, times, nt, nus, nnu);
    if (!d_tn) return KS_ERR_HIP;
    int rc = launch_transfer(c, m, d_tn, nt, nnu, L, nmax, slab, nullptr, nullptr);
    if (rc) return rc;
    const int64_t items = c->npix * frames;
    HIPCHK(launch_seg_combine(x->stream, c->soa, c->npix, c->N, c->M, has_samples(c) ? c->chunk : 0, nt, d_tn + nt, nnu, c->seg, c->out));
    return KS_OK;
}

int ks_cache_march_precision(ks_cache* c, int precision) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (precision != KS_PRECISION_FP64 && precision != KS_PRECISION_MIXED) return fail(KS_ERR_ARG, "unknown precision policy %d", precision);
    c->march_precision = precision;
    return KS_OK;
}

int ks_cache_precision(ks_cache* c, int precision) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (precision != KS_PRECISION_FP64 && precision != KS_PRECISION_MIXED) return fail(KS_ERR_ARG, "unknown precision policy %d", precision);
    c->precision = precision;
    return KS_OK;
}

static int ensure_adj(ks_cache* c, int64_t items, int ns);
static int reverse_launches(ks_cache* c, const ks_model* m, const double* d_tn, int nt, int nnu);

int ks_cache_tuning(ks_cache* c, int chunk, int queue, int patch_a, int patch_b) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (chunk < 8 || chunk > 1024) return fail(KS_ERR_ARG, "samples per work item %d outside 8:1024", chunk);
    if (queue < 0 || queue > (KS_QUEUE_TRANSFER | KS_QUEUE_ADJOINT))
        return fail(KS_ERR_ARG, "queue takes the bits KS_QUEUE_TRANSFER and KS_QUEUE_ADJOINT (got %d)", queue);
    if (patch_a < 1 || patch_b < 1 || patch_a * patch_b > 4096) return fail(KS_ERR_ARG, "wavefront patch %d x %d outside 1:4096 points", patch_a, patch_b);
    HIPCHK(hipSetDevice(c->ctx->device));
    c->fw_model = 0;                   // segment maps of another chunk length would not match a later reverse pass
    c->chunk = chunk; c->queue = queue;
    if (patch_a != c->patch_a || patch_b != c->patch_b) {
        c->patch_a = patch_a; c->patch_b = patch_b;
        if (c->generation > 0) return build_tmap(c);
    }
    return KS_OK;
}

int ks_cache_get_tuning(const ks_cache* c, int* chunk, int* queue, int* patch_a, int* patch_b) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (chunk) *chunk = c->chunk;
    if (queue) *queue = c->queue;
    if (patch_a) *patch_a = c->patch_a;
    if (patch_b) *patch_b = c->patch_b;
    return KS_OK;
}

int ks_cache_set_parts(ks_cache* c, int nparts, const int* n_alpha, const int* n_beta) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (nparts < 0 || (nparts > 0 && (!n_alpha || !n_beta))) return fail(KS_ERR_ARG, "ks_cache_set_parts: null part sizes");
    std::vector<int64_t> off;
    int64_t tot = 0;
    for (int q = 0; q < nparts; ++q) {
        if (n_alpha[q] < 1 || n_beta[q] < 1) return fail(KS_ERR_ARG, "part %d has size %d x %d", q, n_alpha[q], n_beta[q]);
        off.push_back(tot);
        tot += (int64_t)n_alpha[q] * n_beta[q];
    }
    if (nparts > 0 && tot != c->npix) return fail(KS_ERR_ARG, "the parts hold %lld points, the cache %lld", (long long)tot, (long long)c->npix);
    HIPCHK(hipSetDevice(c->ctx->device));
    c->part_off = off;
    c->part_na.assign(n_alpha, n_alpha + nparts);
    c->part_nb.assign(n_beta, n_beta + nparts);
    return c->generation > 0 ? build_tmap(c) : KS_OK;
}

int ks_queue_costs(const ks_cache* c, int64_t* ntasks, uint32_t* cost, int cap) {
    if (!c || !ntasks) return fail(KS_ERR_ARG, "null argument");
    if (c->q_shape[0] < 0 || !c->qcost.p) return fail(KS_ERR_STATE, "no work-queue launch has run on this cache");
    *ntasks = c->q_shape[0];
    if (cost) {
        HIPCHK(hipStreamSynchronize(c->ctx->stream));
        HIPCHK(hipMemcpy(cost, c->qcost.p, (size_t)std::min<int64_t>(cap, c->q_shape[0]) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    return KS_OK;
}

// The autotuner. Transfer (unless KS_TUNE_KEEP_TRANSFER): every candidate launch (samples per work item x wavefront
// patch x launch form) timed with HIP events on a sample of the transfer's work (each task of the candidate's own
// task list with probability 1 / sample_every, by a hash of its index: the same fraction of the (item, sample)
// pairs for every candidate, spread over the whole screen), twice, the faster run counting. Element adjoints (with
// KS_TUNE_ADJOINT): one forward pass with the records at the chosen transfer, then the reverse pass once in each
// launch form of the element adjoints, timed by their stage events (zero seeds: the work does not depend on them).
// The cache keeps the fastest of each; rows of the table: (chunk, patch_a, patch_b, queue bits, ms), the element
// adjoints' rows with chunk = 0.
int ks_autotune(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int flags, int sample_every,
                double* table, int cap, int* ncand) {
    if (!c || !m || !times || !nus || !ncand) return fail(KS_ERR_ARG, "ks_autotune: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (!has_samples(c)) return fail(KS_ERR_ARG, "the autotuner tunes the transfer over stored samples: build the cache with store_samples = true");
    if (m->kind == KS_MODEL_THIN) return fail(KS_ERR_ARG, "thin splats render through ks_thin_image");
    if (nt < 1 || nnu < 1 || !(L > 0.0)) return fail(KS_ERR_ARG, "need times, frequencies and a positive length unit");
    if (sample_every < 1) return fail(KS_ERR_ARG, "sample_every must be positive");
    if (flags & ~(KS_TUNE_RECORDS | KS_TUNE_ADJOINT | KS_TUNE_KEEP_TRANSFER)) return fail(KS_ERR_ARG, "unknown autotuner flags %d", flags);
    const bool records = (flags & (KS_TUNE_RECORDS | KS_TUNE_ADJOINT)) != 0;
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    c->fw_model = 0;
    const int64_t frames = (int64_t)nt * nnu, items = c->npix * frames;
    const int chunks[4] = {16, 32, 64, 128};
    const int patches[3][2] = {{8, 8}, {16, 4}, {4, 16}};
    const int adj_bit = c->queue & KS_QUEUE_ADJOINT;
    const double* d_tn = stage_times(c, times, nt, nus, nnu);
    if (!d_tn) return KS_ERR_HIP;
    int n = 0;
    std::vector<double> rows;
    if (!(flags & KS_TUNE_KEEP_TRANSFER)) {
        const int old[4] = {c->chunk, c->queue, c->patch_a, c->patch_b};
        int best = -1;
        double best_ms = INFINITY;
        std::vector<int> tasks;
        Buf<int> dtasks;
        for (int pc = 0; pc < 3; ++pc) {
            c->patch_a = patches[pc][0]; c->patch_b = patches[pc][1];
            if (int rc = build_tmap(c)) return rc;
            for (int ci = 0; ci < 4; ++ci) {
                if (chunks[ci] > std::max(c->N, 16)) continue;
                c->chunk = chunks[ci];
                if (records && ensure_adj(c, items, std::max(m->ns, 1))) return KS_ERR_HIP;
                const int nseg = seg_count(c, c->chunk);
                if (c->seg_cap < (int64_t)SEG_FIELDS * nseg * items) {
                    dfree(c->seg);
                    if (dalloc(&c->seg, (size_t)SEG_FIELDS * nseg * items)) return KS_ERR_HIP;
                    c->seg_cap = (int64_t)SEG_FIELDS * nseg * items;
                }
                const int64_t all = transfer_tasks(c->npix, nt, nnu, c->N, c->chunk);
                // a hashed Bernoulli sample of rate 1 / sample_every (a regular stride aliases with the screen: every k-th
                // task can fall on one column of patches)
                tasks.clear();
                for (int64_t q = 0; q < all; ++q) {
                    uint64_t h = (uint64_t)q * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
                    h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 29;
                    if (h % (uint64_t)sample_every == 0) tasks.push_back((int)q);
                }
                if (tasks.empty()) tasks.push_back(0);
                if (dtasks.grow(tasks.size())) return KS_ERR_HIP;
                HIPCHK(hipMemcpy(dtasks.p, tasks.data(), tasks.size() * sizeof(int), hipMemcpyHostToDevice));
                TransferLaunch smp;
                smp.tasks = dtasks.p; smp.ntasks = (int64_t)tasks.size();
                for (int qu = 0; qu < 2; ++qu) {
                    c->queue = qu | adj_bit;
                    float ms = INFINITY;
                    for (int rep = 0; rep < 2; ++rep) {
                        HIPCHK(hipEventRecord(x->ev0, x->stream));
                        int rc = transfer_launch(c, m, d_tn, nt, nnu, L, nullptr, records ? (void*)c->arec : nullptr, records ? c->aflg : nullptr, &smp);
                        if (rc) { dtasks.release(); return rc; }
                        HIPCHK(hipEventRecord(x->ev1, x->stream));
                        HIPCHK(hipEventSynchronize(x->ev1));
                        float t = 0.0f;
                        HIPCHK(hipEventElapsedTime(&t, x->ev0, x->ev1));
                        ms = std::min(ms, t);
                    }
                    rows.insert(rows.end(), {(double)c->chunk, (double)c->patch_a, (double)c->patch_b, (double)c->queue, (double)ms});
                    if (ms < best_ms) { best_ms = ms; best = n; }
                    ++n;
                }
            }
        }
        dtasks.release();
        if (best < 0) { c->chunk = old[0]; c->queue = old[1]; c->patch_a = old[2]; c->patch_b = old[3]; }
        else {
            c->chunk = (int)rows[5 * best]; c->patch_a = (int)rows[5 * best + 1]; c->patch_b = (int)rows[5 * best + 2];
            c->queue = (int)rows[5 * best + 3];
        }
        if (int rc = build_tmap(c)) return rc;
    }
    if (flags & KS_TUNE_ADJOINT) {
        if (m->ns < 1) return fail(KS_ERR_ARG, "the element adjoints need splats");
        if (ensure_out(c, frames) || ensure_adj(c, items, m->ns)) return KS_ERR_HIP;
        if (int rc = launch_transfer(c, m, d_tn, nt, nnu, L, -1, 0.0, c->arec, c->aflg)) return rc;
        c->fw_chunk = c->chunk;
        HIPCHK(hipMemsetAsync(c->ads, 0, (size_t)4 * items * sizeof(double), x->stream));
        int best_q = c->queue;
        double best_ms = INFINITY;
        for (int qa = 0; qa < 2; ++qa) {
            c->queue = (c->queue & KS_QUEUE_TRANSFER) | (qa ? KS_QUEUE_ADJOINT : 0);
            if (int rc = reverse_launches(c, m, d_tn, nt, nnu)) return rc;
            HIPCHK(hipStreamSynchronize(x->stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, x->evs[KS_STAGE_ADJ_ELEMENTS][0], x->evs[KS_STAGE_ADJ_ELEMENTS][1]));
            rows.insert(rows.end(), {0.0, 0.0, 0.0, (double)c->queue, (double)ms});
            ++n;
            if (ms < best_ms) { best_ms = ms; best_q = c->queue; }
        }
        c->queue = best_q;
    }
    c->fw_model = 0;
    *ncand = n;
    if (table) std::copy(rows.begin(), rows.begin() + 5 * std::min(n, std::max(cap, 0)), table);
    return KS_OK;
}

// re-run the cache's marcher into its sample storage on the current pixel constants
int ks_march(ks_cache* c) {
    if (!c) return fail(KS_ERR_ARG, "null cache");
    if (!has_samples(c)) return fail(KS_ERR_STATE, "the cache holds no stored samples: build it with store_samples = true and a storing marcher");
    HIPCHK(hipSetDevice(c->ctx->device));
    return launch_store_march(c);
}

int ks_render(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
              double* out_stokes, int screen_order) {
    int rc = ks_render_device(c, m, times, nt, nus, nnu, L, nmax, slab);
    if (rc) return rc;
    return ks_copy_stokes(c, (int64_t)nt * nnu, out_stokes, screen_order);
}

// ---- the dual-sweep gradient (polarized_gradient!, polarized.jl:473-480; dual_gradient.jl:180-249) ----
// Units per launch of the element adjoints: their partial sums (units x splats x ADJ_ROWS doubles) stay within
// ADJ_PART_MAX doubles (4 GiB), in whole groups of 8 units (the XCD order of the launch).
constexpr int64_t ADJ_PART_MAX = (int64_t)1 << 29;
static int64_t adj_unit_batch(int64_t nunits, int ns) {
    const int64_t per = (int64_t)std::max(ns, 1) * ADJ_ROWS;
    int64_t nb = std::max<int64_t>(8, ADJ_PART_MAX / per / 8 * 8);
    return std::min(nb, (nunits + 7) / 8 * 8);
}

static int ensure_adj(ks_cache* c, int64_t items, int ns) {
    const int nseg = (c->N + c->chunk - 1) / c->chunk;
    if (c->a_items < items) {
        dfree(c->arec); dfree(c->abar); dfree(c->aflg); dfree(c->atail); dfree(c->ads);
        c->a_items = 0;
        const size_t per = (size_t)c->N * items;
        if (dalloc(&c->arec, (size_t)REC_FIELDS * per) || dalloc(&c->abar, (size_t)BAR_FIELDS * per) || dalloc(&c->aflg, per) || dalloc(&c->atail, 4 * per) ||
            dalloc(&c->ads, (size_t)4 * items))
            return KS_ERR_HIP;
        c->a_items = items;
    }
    if (c->a_wT < (int64_t)8 * nseg * items) {      // per chunk (w_in, T_out): grows with the items and the chunks per ray
        dfree(c->awT);
        c->a_wT = 0;
        if (dalloc(&c->awT, (size_t)8 * nseg * items)) return KS_ERR_HIP;
        c->a_wT = (int64_t)8 * nseg * items;
    }
    const int64_t nunits = adj_units(items, c->N);
    const int64_t nb = adj_unit_batch(nunits, ns);
    if (c->a_part < nb * ns * ADJ_ROWS) {
        dfree(c->apart);
        if (dalloc(&c->apart, (size_t)nb * ns * ADJ_ROWS)) return KS_ERR_HIP;
        c->a_part = nb * ns * ADJ_ROWS;
    }
    if (c->a_ub < nunits) {
        dfree(c->aub);
        if (dalloc(&c->aub, (size_t)nunits * UB_FIELDS)) return KS_ERR_HIP;
        c->a_ub = nunits;
    }
    if (c->a_dpar < (int64_t)ns * ADJ_ROWS) {
        dfree(c->adpar);
        if (dalloc(&c->adpar, (size_t)ns * ADJ_ROWS)) return KS_ERR_HIP;
        c->a_dpar = (int64_t)ns * ADJ_ROWS;
    }
    return KS_OK;
}

// the reverse pass's launches after the seeds: tails and adjoints per chunk, the element adjoints in batches, the
// reduction into c->adpar (ns x ADJ_ROWS)
static int reverse_launches(ks_cache* c, const ks_model* m, const double* d_tn, int nt, int nnu) {
    ks_ctx* x = c->ctx;
    const int64_t items = c->npix * nt * nnu;
    const int ns = m->ns;
    HIPCHK(launch_adj_combine(x->stream, c->npix, c->N, c->fw_chunk, nt, d_tn + nt, nnu, c->seg, c->ads, c->awT, nullptr));
    HIPCHK(x->stage_begin(KS_STAGE_ADJ_STEPS));
    const bool mixed = c->precision == KS_PRECISION_MIXED;
    if (mixed) HIPCHK(launch_adj_segment_f32(x->stream, c->N, c->fw_chunk, items, c->awT, (const f32*)c->arec, (f32*)c->abar, c->aflg, (f32*)c->atail,
                                             samples_of(c), c->npix));
    else HIPCHK(launch_adj_segment(x->stream, c->N, c->fw_chunk, items, c->awT, c->arec, c->abar, c->aflg, c->atail, samples_of(c), c->npix));
    HIPCHK(x->stage_end(KS_STAGE_ADJ_STEPS));
    if (ns > 0) {
        if (mixed)
            if (int rc = ensure_p4(c)) return rc;
        HIPCHK(x->stage_begin(KS_STAGE_ADJ_ELEMENTS));
        HIPCHK(launch_unit_bounds(x->stream, c->npix, c->N, samples_of(c), d_tn, nt, items, c->aflg, c->aub));
        // the persistent form of the element adjoints with the cache's work-queue choice: per-XCD queue heads
        int* heads = nullptr;
        if (c->queue & KS_QUEUE_ADJOINT) {
            if (c->aheads.grow(8)) return KS_ERR_HIP;
            heads = c->aheads.p;
        }
        // the units in batches whose partial sums fit the workspace, each reduced into the running sums in order
        const int64_t nunits = adj_units(items, c->N), nbmax = adj_unit_batch(nunits, ns);
        for (int64_t u0 = 0; u0 < nunits; u0 += nbmax) {
            const int64_t nb = std::min(nbmax, nunits - u0);
            if (mixed)
                HIPCHK(launch_adj_splats_f32(x->stream, m->pops, c->spin, c->thetao, c->soa, c->npix, c->N, samples_of(c), m->d_splats_f, m->d_adj_f, ns,
                                             m->d_knots_f, m->d_ktimes_f, d_tn, nt, d_tn + nt, items, (const f32*)c->arec, (const f32*)c->abar, c->aflg,
                                             u0, nb, c->aub, c->apart, heads));
            else
                HIPCHK(launch_adj_splats(x->stream, m->pops, c->spin, c->thetao, c->soa, c->npix, c->N, samples_of(c), m->d_splats, m->d_adj, ns, m->d_knots,
                                         m->d_ktimes, d_tn, nt, d_tn + nt, items, c->arec, c->abar, c->aflg, u0, nb, c->aub, c->apart, heads));
            HIPCHK(launch_adj_reduce(x->stream, c->apart, nb, ns, u0 > 0, c->adpar));
        }
        HIPCHK(x->stage_end(KS_STAGE_ADJ_ELEMENTS));
    }
    return KS_OK;
}

// forward pass of the dual sweep: the transfer with the per-sample records, and the images
int ks_gradient_forward(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
                        double* image, int screen_order) {
    if (!c || !m || !times || !nus) return fail(KS_ERR_ARG, "ks_gradient_forward: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (m->kind == KS_MODEL_THIN) return fail(KS_ERR_ARG, "the dual sweep takes polarized splats (thermal, power-law, kappa, knot or composite)");
    if (!has_samples(c)) return fail(KS_ERR_ARG, "the cache holds no stored samples: build it with store_samples = true and a storing marcher");
    if (nt < 1 || nnu < 1) return fail(KS_ERR_ARG, "need at least one time and one frequency");
    if (!(L > 0.0)) return fail(KS_ERR_ARG, "length unit L must be positive");
    for (int i = 0; i < nnu; ++i) if (!(nus[i] > 0.0)) return fail(KS_ERR_ARG, "frequencies must be positive");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    c->fw_model = 0;
    const int64_t frames = (int64_t)nt * nnu;
    const int64_t items = c->npix * frames;
    const int ns = m->ns;
    if (ensure_out(c, frames) || ensure_adj(c, items, ns > 0 ? ns : 1)) return KS_ERR_HIP;
    const double* d_tn = stage_times(c, times, nt, nus, nnu);
    if (!d_tn) return KS_ERR_HIP;
    int rc = launch_transfer(c, m, d_tn, nt, nnu, L, nmax, slab, c->arec, c->aflg);
    if (rc) return rc;
    HIPCHK(launch_seg_combine(x->stream, c->soa, c->npix, c->N, c->M, c->chunk, nt, d_tn + nt, nnu, c->seg, c->out));
    c->fw_model = m->id; c->fw_nt = nt; c->fw_nnu = nnu; c->fw_generation = c->generation; c->fw_precision = c->precision;
    c->fw_chunk = c->chunk;
    c->fw_seeded.assign((size_t)frames, 0);
    if (image) return ks_copy_stokes(c, frames, image, screen_order);
    return KS_OK;
}

// reverse pass: the seeds, the tails and adjoints per chunk, the element adjoints, the gradient
int ks_gradient_reverse(ks_cache* c, const ks_model* m, const double* dstokes, int screen_order, double* dparams) {
    if (!c || !m || !dparams) return fail(KS_ERR_ARG, "ks_gradient_reverse: null argument");
    if (!c->fw_model) return fail(KS_ERR_STATE, "no forward pass to reverse: call ks_gradient_forward first");
    if (m->id != c->fw_model) return fail(KS_ERR_ARG, "the reverse pass takes the model of its forward pass");
    if (c->generation != c->fw_generation || c->precision != c->fw_precision)
        return fail(KS_ERR_STATE, "the cache changed (regenerate! or precision) since the forward pass");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int nt = c->fw_nt, nnu = c->fw_nnu;
    const int64_t frames = (int64_t)nt * nnu;
    const int64_t items = c->npix * frames;
    const int ns = m->ns;
    if (dstokes) {
        // seeds in sorted order: ads[4 (j + npix f) + q]
        std::vector<double> ds((size_t)4 * items);
        for (int64_t f = 0; f < frames; ++f)
            for (int64_t j = 0; j < c->npix; ++j) {
                const int64_t src = screen_order ? c->perm_host[j] + c->npix * f : j + c->npix * f;
                for (int q = 0; q < 4; ++q) ds[4 * (j + c->npix * f) + q] = dstokes[4 * src + q];
            }
        HIPCHK(hipMemcpyAsync(c->ads, ds.data(), ds.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));     // ds is released on every return path below
        std::fill(c->fw_seeded.begin(), c->fw_seeded.end(), 0);     // the device seeds are replaced
    } else {
        for (int64_t f = 0; f < frames; ++f)
            if (!c->fw_seeded[f]) return fail(KS_ERR_STATE, "frame %lld has no device seed: give dstokes or call a *_chi2_seed for every frame", (long long)f);
    }
    if (int rc = reverse_launches(c, m, c->dtn, nt, nnu)) return rc;
    std::vector<double> g((size_t)ns * ADJ_ROWS);
    if (ns > 0) HIPCHK(hipMemcpyAsync(g.data(), c->adpar, g.size() * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    // every splat's rows into its column of dparams (21 rows thermal and knot, 22 power law and kappa; a composite's
    // parts in order, each part nrows x ns column-major)
    size_t off = 0;
    for (int i = 0; i < ns; ++i) {
        const int nr = m->hrows[i];
        for (int r = 0; r < nr; ++r) dparams[off + r] += g[(size_t)i * ADJ_ROWS + r];
        off += nr;
    }
    return KS_OK;
}

int ks_polarized_gradient(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
                          const double* dstokes, int screen_order, double* dparams, double* image) {
    if (!c || !m || !times || !nus || !dstokes || !dparams) return fail(KS_ERR_ARG, "ks_polarized_gradient: null argument");
    int rc = ks_gradient_forward(c, m, times, nt, nus, nnu, L, nmax, slab, image, screen_order);
    if (rc) return rc;
    return ks_gradient_reverse(c, m, dstokes, screen_order, dparams);
}

int ks_copy_stokes(ks_cache* c, int64_t frames, double* out_stokes, int screen_order) {
    if (!c || !out_stokes) return fail(KS_ERR_ARG, "null argument");
    if (frames > c->out_frames) return fail(KS_ERR_DIM, "only %lld frames rendered", (long long)c->out_frames);
    ks_ctx* x = c->ctx;
    size_t n = (size_t)4 * c->npix * frames;
    const double* src = c->out;
    if (screen_order) {
        HIPCHK(launch_unsort4(x->stream, c->out, c->perm, c->npix, frames, c->out_screen));
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
    const int nseg = max_segments(c->N);
    if (!c->thin_part) { if (dalloc(&c->thin_part, (size_t)nseg * c->npix)) return KS_ERR_HIP; }
    // per segment sums from the stored samples (CHUNK samples) or fused with the march (anchor segments)
    HIPCHK(launch_thin_parts(x->stream, c->spin, c->thetao, c->soa, c->npix, c->N, c->M, has_samples(c) ? samples_of(c) : no_samples(),
                             m->d_splats, m->ns, t_obs, nseg, c->thin_part, c->segres));
    HIPCHK(launch_thin_combine(x->stream, c->npix, nseg, c->thin_part, c->img1));
    if (!has_samples(c)) HIPCHK(launch_res_reduce(x->stream, c->npix, nseg, c->segres, c->rt, c->rp));
    const double* src = c->img1;
    if (screen_order) {
        HIPCHK(launch_unsort1(x->stream, c->img1, c->perm, c->npix, c->img1 + c->npix));
        src = c->img1 + c->npix;
    }
    HIPCHK(hipMemcpyAsync(out, src, c->npix * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

// ---- per-ray splat lists (ray_lists, dual_gradient.jl:283-359) ----
int ks_ray_lists(ks_cache* c, const ks_model* m, const double* times, int nt, int nmax, double slab, int screen_order, int32_t* count,
                 int* capacity, int32_t* ids) {
    if (!c || !m || !times || !count || !capacity) return fail(KS_ERR_ARG, "ks_ray_lists: null argument");
    if (nt < 1) return fail(KS_ERR_ARG, "ks_ray_lists: no frames");
    if (!has_samples(c))
        return fail(KS_ERR_STATE, "the cache holds no stored samples: build it with store_samples and a storing marcher");
    if (m->kind != KS_MODEL_THERMAL && m->kind != KS_MODEL_POWERLAW && m->kind != KS_MODEL_KAPPA)
        return fail(KS_ERR_ARG, "ks_ray_lists takes thermal, power-law or kappa splats");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int64_t npix = c->npix, items = npix * nt;
    struct Scratch {   // freed on every return
        Buf<uint8_t> hits; Buf<int> cnt; Buf<double> dt; Buf<int32_t> ids;
        ~Scratch() { hits.release(); cnt.release(); dt.release(); ids.release(); }
    } b;
    Buf<uint8_t>& hits = b.hits;
    Buf<int>& cnt = b.cnt;
    Buf<double>& dt = b.dt;
    Buf<int32_t>& dids = b.ids;
    if (hits.grow((size_t)std::max(m->ns, 1) * items) || cnt.grow(items) || dt.grow(nt)) return KS_ERR_HIP;
    HIPCHK(hipMemcpyAsync(dt.p, times, nt * sizeof(double), hipMemcpyHostToDevice, x->stream));
    const int* kst = nullptr;
    if (nmax >= 0) {
        if (!c->kstop) { if (dalloc(&c->kstop, (size_t)npix)) return KS_ERR_HIP; }
        HIPCHK(launch_wind_cutoff(x->stream, c->spin, c->thetao, c->soa, npix, c->N, c->M, nmax, slab, samples_of(c), c->kstop));
        kst = c->kstop;
    }
    const double rp = 1.0 + sqrt(1.0 - c->spin * c->spin);
    HIPCHK(launch_ray_hits(x->stream, rp, npix, c->N, samples_of(c), m->d_splats, m->ns, dt.p, nt, kst, hits.p, cnt.p));
    std::vector<int> cs(items);
    HIPCHK(hipMemcpyAsync(cs.data(), cnt.p, items * sizeof(int), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    int cap = 1;
    for (int64_t t = 0; t < items; ++t) {
        cap = std::max(cap, cs[t]);
        const int64_t j = t % npix, f = t / npix;
        count[(screen_order ? c->perm_host[j] : j) + npix * f] = cs[t];
    }
    if (ids) {
        if (*capacity < cap) {
            const int have = *capacity;
            *capacity = cap;
            return fail(KS_ERR_DIM, "ks_ray_lists: capacity %d below the largest count %d", have, cap);
        }
        const int cap_out = *capacity;
        if (dids.grow((size_t)cap_out * items)) return KS_ERR_HIP;
        HIPCHK(launch_ray_fill(x->stream, items, npix, m->ns, cap_out, hits.p, screen_order ? c->iperm : nullptr, dids.p));
        HIPCHK(hipMemcpyAsync(ids, dids.p, (size_t)cap_out * items * sizeof(int32_t), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    } else {
        *capacity = cap;
    }
    return KS_OK;
}

// ---- visibilities and χ² (visibilities.jl:21-40, device_visibilities.jl:53-111, Fit.jl:537-577) ----
// device workspace of the visibility transform: (u, v, taper), the phase tables A2 (2 nuv x n_alpha) and
// B2 (2 nuv x n_beta), the frames as n_alpha x 4 n_beta matrices, the partial products C2 (2 nuv x 4 n_beta per
// frame) and V (nframes x nuv x 4 complex)
struct VisWs { double *u, *v, *t, *A2, *B2, *Im, *C2, *V; };

// host (u, v, taper) to the device and the phase tables of the cache's screen
static int vis_tables_ws(ks_cache* c, const VisWs& w, double psize, const double* u, const double* v, const double* taper, int nuv) {
    ks_ctx* x = c->ctx;
    HIPCHK(hipMemcpyAsync(w.u, u, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(w.v, v, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    if (taper) HIPCHK(hipMemcpyAsync(w.t, taper, nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(launch_vis_tables(x->stream, nuv, c->nalpha, c->nbeta, psize, w.u, w.v, w.A2, w.B2));
    return KS_OK;
}

// V of frames [f0, f0 + nf) of the images in the cache into w.V
static int vis_forward_ws(ks_cache* c, const VisWs& w, int64_t f0, int nf, int nuv, const double* d_tap, double scale) {
    ks_ctx* x = c->ctx;
    const int na = c->nalpha, nb = c->nbeta;
    const int64_t mat = (int64_t)na * 4 * nb, cmat = (int64_t)2 * nuv * 4 * nb;
    HIPCHK(launch_vis_gather(x->stream, c->out, c->perm, c->npix, na, nb, f0, nf, w.Im));
    HIPCHK(x->stage_begin(KS_STAGE_VIS));
    HIPCHK(launch_dgemm(x->stream, false, 2 * nuv, 4 * nb, na, w.A2, na, 0, w.Im, 4 * nb, mat, w.C2, 4 * nb, cmat, nf));
    HIPCHK(x->stage_end(KS_STAGE_VIS));
    HIPCHK(launch_vis_combine(x->stream, nuv, nb, nf, w.C2, w.B2, d_tap, scale, w.V));
    return KS_OK;
}

static VisWs cache_ws(const ks_cache* c) { return VisWs{c->vu.p, c->vv.p, c->vt.p, c->vA2.p, c->vB2.p, c->vIm.p, c->vC2.p, c->vV.p}; }

// the phase tables in the cache's grow-only workspace
static int vis_setup(ks_cache* c, double psize, const double* u, const double* v, const double* taper, int nuv) {
    const int na = c->nalpha, nb = c->nbeta;
    if (c->vu.grow(nuv) || c->vv.grow(nuv) || c->vt.grow(nuv) || c->vA2.grow((size_t)2 * nuv * na) || c->vB2.grow((size_t)2 * nuv * nb))
        return KS_ERR_HIP;
    return vis_tables_ws(c, cache_ws(c), psize, u, v, taper, nuv);
}

// V of frames [f0, f0 + nf) into c->vV
static int vis_forward(ks_cache* c, int64_t f0, int nf, int nuv, const double* d_tap, double scale) {
    const int na = c->nalpha, nb = c->nbeta;
    const int64_t mat = (int64_t)na * 4 * nb, cmat = (int64_t)2 * nuv * 4 * nb;
    if (c->vIm.grow((size_t)mat * nf) || c->vC2.grow((size_t)cmat * nf) || c->vV.grow((size_t)8 * nuv * nf)) return KS_ERR_HIP;
    return vis_forward_ws(c, cache_ws(c), f0, nf, nuv, d_tap, scale);
}

static int screen_ready(const ks_cache* c) {
    if ((int64_t)c->nalpha * c->nbeta != c->npix) return fail(KS_ERR_STATE, "the cache's screen shape is not set");
    return KS_OK;
}

// visibilities of the last rendered frames (images in the cache, sorted order), EHT sign convention
int ks_visibilities(ks_cache* c, int64_t frames, double psize, const double* u, const double* v, const double* taper, int nuv, double* V) {
    if (!c || !u || !v || !V) return fail(KS_ERR_ARG, "ks_visibilities: null argument");
    if (frames > c->out_frames || frames < 1) return fail(KS_ERR_DIM, "only %lld frames rendered", (long long)c->out_frames);
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if (int rc = screen_ready(c)) return rc;
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    if (int rc = vis_setup(c, psize, u, v, taper, nuv)) return rc;
    const double scale = psize * psize / JY;
    if (int rc = vis_forward(c, 0, (int)frames, nuv, taper ? c->vt.p : nullptr, scale)) return rc;
    HIPCHK(hipMemcpyAsync(V, c->vV.p, (size_t)8 * nuv * frames * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

// ks_visibilities without allocating: the same transform with its workspace carved from the device array
// `scratch` (KS_NOALLOC_SCRATCH_BYTES) at the address recorded for the context's device; each part starts on a
// 256-byte boundary. Needs nothing cleared beforehand: every kernel of the transform writes its whole output.
int ks_visibilities_noalloc(ks_cache* c, int64_t frames, double psize, const double* u, const double* v, const double* taper, int nuv, double* V) {
    if (!c || !u || !v || !V) return fail(KS_ERR_ARG, "ks_visibilities_noalloc: null argument");
    if (frames > c->out_frames || frames < 1) return fail(KS_ERR_DIM, "only %lld frames rendered", (long long)c->out_frames);
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if (int rc = screen_ready(c)) return rc;
    ks_ctx* x = c->ctx;
    char* base = (char*)x->scratch_dev;
    if (!base) return fail(KS_ERR_STATE, "the context's device has no scratch address");
    const int na = c->nalpha, nb = c->nbeta;
    const int64_t parts[8] = {nuv, nuv, nuv, (int64_t)2 * nuv * na, (int64_t)2 * nuv * nb, (int64_t)na * 4 * nb * frames,
                              (int64_t)2 * nuv * 4 * nb * frames, (int64_t)8 * nuv * frames};
    double* ptr[8];
    int64_t off = 0;
    for (int k = 0; k < 8; ++k) {
        ptr[k] = (double*)(base + off);
        off += (parts[k] * (int64_t)sizeof(double) + 255) / 256 * 256;
    }
    if (off > (int64_t)KS_NOALLOC_SCRATCH_BYTES)
        return fail(KS_ERR_DIM, "ks_visibilities_noalloc: %lld frames of a %d x %d screen at %d (u, v) points need %lld bytes of workspace, "
                    "above the fixed %lld", (long long)frames, na, nb, nuv, (long long)off, (long long)KS_NOALLOC_SCRATCH_BYTES);
    const VisWs w{ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], ptr[6], ptr[7]};
    HIPCHK(hipSetDevice(x->device));
    if (int rc = vis_tables_ws(c, w, psize, u, v, taper, nuv)) return rc;
    if (int rc = vis_forward_ws(c, w, 0, (int)frames, nuv, taper ? w.t : nullptr, psize * psize / JY)) return rc;
    HIPCHK(hipMemcpyAsync(V, w.V, (size_t)8 * nuv * frames * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return KS_OK;
}

static int seed_frame_ok(ks_cache* c, int64_t frame) {
    if (!c->fw_model) return fail(KS_ERR_STATE, "no forward pass: call ks_gradient_forward first");
    if (frame < 0 || frame >= (int64_t)c->fw_seeded.size())
        return fail(KS_ERR_DIM, "frame %lld outside the forward pass's %zu frames", (long long)frame, c->fw_seeded.size());
    return KS_OK;
}

// frame_chi2_seed!: χ² of one frame's visibilities and the dual sweep's seed of that frame on the device
int ks_vis_chi2_seed(ks_cache* c, int64_t frame, double psize, const double* u, const double* v, const double* taper, const double* gains,
                     const double* vis, const double* sigma, int nuv, double* chi2, double* V) {
    if (!c || !u || !v || !vis || !sigma || !chi2) return fail(KS_ERR_ARG, "ks_vis_chi2_seed: null argument");
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if (int rc = seed_frame_ok(c, frame)) return rc;
    if (int rc = screen_ready(c)) return rc;
    for (int k = 0; k < 4 * nuv; ++k) if (!(sigma[k] > 0.0)) return fail(KS_ERR_ARG, "sigma must be positive (entry %d)", k);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int na = c->nalpha, nb = c->nbeta;
    if (int rc = vis_setup(c, psize, u, v, taper, nuv)) return rc;
    if (c->vg.grow((size_t)2 * nuv) || c->vd.grow((size_t)8 * nuv) || c->vs.grow((size_t)4 * nuv) || c->vwq.grow((size_t)8 * nuv) ||
        c->vE2.grow((size_t)2 * nuv * 4 * nb) || c->vSd.grow((size_t)na * 4 * nb) || c->vSp.grow((size_t)VIS_SPLITK * na * 4 * nb) ||
        c->vchi.grow(nuv) || c->vsum.grow(1))
        return KS_ERR_HIP;
    if (gains) HIPCHK(hipMemcpyAsync(c->vg.p, gains, (size_t)2 * nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->vd.p, vis, (size_t)8 * nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->vs.p, sigma, (size_t)4 * nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    const double scale = psize * psize / JY;
    const double* d_tap = taper ? c->vt.p : nullptr;
    if (int rc = vis_forward(c, frame, 1, nuv, d_tap, scale)) return rc;
    HIPCHK(launch_vis_residual(x->stream, nuv, c->vV.p, c->vd.p, c->vs.p, gains ? c->vg.p : nullptr, d_tap, scale, c->vwq.p, c->vchi.p));
    HIPCHK(launch_vis_rotate(x->stream, nuv, nb, c->vwq.p, c->vB2.p, c->vE2.p));
    HIPCHK(launch_dgemm_tn_splitk(x->stream, na, 4 * nb, 2 * nuv, c->vA2.p, na, c->vE2.p, 4 * nb, c->vSd.p, VIS_SPLITK, c->vSp.p));
    HIPCHK(launch_vis_scatter_seed(x->stream, c->vSd.p, c->perm, c->npix, na, nb, frame, 0, c->ads));
    HIPCHK(launch_sum_fixed(x->stream, c->vchi.p, nuv, c->vsum.p));
    HIPCHK(hipMemcpyAsync(chi2, c->vsum.p, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    if (V) HIPCHK(hipMemcpyAsync(V, c->vV.p, (size_t)8 * nuv * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    c->fw_seeded[frame] = 1;
    return KS_OK;
}

// seed_kernel! for an arbitrary visibility-domain loss: the adjoint transform of w = dloss/dV into the
// device seed of one frame of the last forward pass (overwriting it, or added to it with add)
int ks_vis_seed(ks_cache* c, int64_t frame, double psize, const double* u, const double* v, const double* taper, const double* w, int nuv,
                int add) {
    if (!c || !u || !v || !w) return fail(KS_ERR_ARG, "ks_vis_seed: null argument");
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if (int rc = seed_frame_ok(c, frame)) return rc;
    if (int rc = screen_ready(c)) return rc;
    if (add && !c->fw_seeded[frame]) return fail(KS_ERR_STATE, "frame %lld has no seed to add to", (long long)frame);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int na = c->nalpha, nb = c->nbeta;
    if (int rc = vis_setup(c, psize, u, v, taper, nuv)) return rc;
    if (c->vd.grow((size_t)8 * nuv) || c->vwq.grow((size_t)8 * nuv) || c->vE2.grow((size_t)2 * nuv * 4 * nb) || c->vSd.grow((size_t)na * 4 * nb) ||
        c->vSp.grow((size_t)VIS_SPLITK * na * 4 * nb))
        return KS_ERR_HIP;
    HIPCHK(hipMemcpyAsync(c->vd.p, w, (size_t)8 * nuv * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(launch_vis_scale_weights(x->stream, nuv, c->vd.p, taper ? c->vt.p : nullptr, psize * psize / JY, c->vwq.p));
    HIPCHK(launch_vis_rotate(x->stream, nuv, nb, c->vwq.p, c->vB2.p, c->vE2.p));
    HIPCHK(launch_dgemm_tn_splitk(x->stream, na, 4 * nb, 2 * nuv, c->vA2.p, na, c->vE2.p, 4 * nb, c->vSd.p, VIS_SPLITK, c->vSp.p));
    HIPCHK(launch_vis_scatter_seed(x->stream, c->vSd.p, c->perm, c->npix, na, nb, frame, add, c->ads));
    HIPCHK(hipStreamSynchronize(x->stream));
    c->fw_seeded[frame] = 1;
    return KS_OK;
}

int ks_closure_chi2_seed(ks_cache* c, int64_t frame, double psize, const double* u, const double* v, const double* taper, int nuv,
                         int ntri, const int* triangles, const double* phases, const double* sigma_phase, int nquad, const int* quadrangles,
                         const double* logamps, const double* sigma_logamp, int add, double* chi2) {
    if (!c || !u || !v || !chi2) return fail(KS_ERR_ARG, "ks_closure_chi2_seed: null argument");
    if (nuv < 1) return fail(KS_ERR_ARG, "need at least one (u, v) point");
    if (ntri < 0 || nquad < 0 || ntri + nquad < 1) return fail(KS_ERR_ARG, "need at least one closure triangle or quadrangle");
    if ((ntri > 0 && (!triangles || !phases || !sigma_phase)) || (nquad > 0 && (!quadrangles || !logamps || !sigma_logamp)))
        return fail(KS_ERR_ARG, "ks_closure_chi2_seed: null closure data");
    if (int rc = seed_frame_ok(c, frame)) return rc;
    if (int rc = screen_ready(c)) return rc;
    if (add && !c->fw_seeded[frame]) return fail(KS_ERR_STATE, "frame %lld has no seed to add to", (long long)frame);
    // the per-baseline lists of the terms that use each baseline, in term order (the legs' order within a term)
    std::vector<int> off(nuv + 1, 0), legs;
    std::vector<std::pair<int, int>> uses;               // (baseline, entry)
    for (int t = 0; t < ntri; ++t) {
        if (!(sigma_phase[t] > 0.0)) return fail(KS_ERR_ARG, "closure phase sigma must be positive (triangle %d)", t);
        for (int l = 0; l < 3; ++l) {
            const int k = triangles[3 * t + l];
            if (k == 0 || k > nuv || k < -nuv) return fail(KS_ERR_ARG, "triangle %d: baseline index %d outside 1:%d", t, k, nuv);
            uses.push_back({std::abs(k) - 1, t | (k < 0 ? 0x40000000 : 0)});
        }
    }
    for (int q = 0; q < nquad; ++q) {
        if (!(sigma_logamp[q] > 0.0)) return fail(KS_ERR_ARG, "log closure amplitude sigma must be positive (quadrangle %d)", q);
        for (int l = 0; l < 4; ++l) {
            const int k = quadrangles[4 * q + l];
            if (k == 0 || k > nuv || k < -nuv) return fail(KS_ERR_ARG, "quadrangle %d: baseline index %d outside 1:%d", q, k, nuv);
            uses.push_back({std::abs(k) - 1, (ntri + q) | (l >= 2 ? 0x40000000 : 0)});
        }
    }
    std::stable_sort(uses.begin(), uses.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return x.first < y.first; });
    legs.reserve(uses.size());
    for (const auto& e : uses) { off[e.first + 1] += 1; legs.push_back(e.second); }
    for (int k = 0; k < nuv; ++k) off[k + 1] += off[k];
    std::vector<double> dat;
    dat.insert(dat.end(), phases, phases + ntri); dat.insert(dat.end(), sigma_phase, sigma_phase + ntri);
    dat.insert(dat.end(), logamps, logamps + nquad); dat.insert(dat.end(), sigma_logamp, sigma_logamp + nquad);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int na = c->nalpha, nb = c->nbeta, nt = ntri + nquad;
    if (int rc = vis_setup(c, psize, u, v, taper, nuv)) return rc;
    if (c->ctri.grow((size_t)std::max(3 * ntri, 1)) || c->cquad.grow((size_t)std::max(4 * nquad, 1)) || c->coff.grow(off.size()) ||
        c->clegs.grow(std::max(legs.size(), (size_t)1)) || c->cdat.grow(std::max(dat.size(), (size_t)1)) || c->cterm.grow((size_t)2 * nt) ||
        c->vd.grow((size_t)8 * nuv) || c->vwq.grow((size_t)8 * nuv) || c->vE2.grow((size_t)2 * nuv * 4 * nb) || c->vSd.grow((size_t)na * 4 * nb) ||
        c->vSp.grow((size_t)VIS_SPLITK * na * 4 * nb) || c->vsum.grow(1))
        return KS_ERR_HIP;
    if (ntri) HIPCHK(hipMemcpyAsync(c->ctri.p, triangles, (size_t)3 * ntri * sizeof(int), hipMemcpyHostToDevice, x->stream));
    if (nquad) HIPCHK(hipMemcpyAsync(c->cquad.p, quadrangles, (size_t)4 * nquad * sizeof(int), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->coff.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->clegs.p, legs.data(), legs.size() * sizeof(int), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->cdat.p, dat.data(), dat.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
    const double scale = psize * psize / JY;
    const double* d_tap = taper ? c->vt.p : nullptr;
    if (int rc = vis_forward(c, frame, 1, nuv, d_tap, scale)) return rc;
    const double* d_cp = c->cdat.p;
    HIPCHK(launch_closure_terms(x->stream, ntri, c->ctri.p, d_cp, d_cp + ntri, nquad, c->cquad.p, d_cp + 2 * ntri, d_cp + 2 * ntri + nquad, c->vV.p,
                                c->cterm.p, c->cterm.p + nt));
    HIPCHK(launch_closure_weights(x->stream, nuv, ntri, c->coff.p, c->clegs.p, c->cterm.p + nt, c->vV.p, c->vd.p));
    HIPCHK(launch_vis_scale_weights(x->stream, nuv, c->vd.p, d_tap, scale, c->vwq.p));
    HIPCHK(launch_vis_rotate(x->stream, nuv, nb, c->vwq.p, c->vB2.p, c->vE2.p));
    HIPCHK(launch_dgemm_tn_splitk(x->stream, na, 4 * nb, 2 * nuv, c->vA2.p, na, c->vE2.p, 4 * nb, c->vSd.p, VIS_SPLITK, c->vSp.p));
    HIPCHK(launch_vis_scatter_seed(x->stream, c->vSd.p, c->perm, c->npix, na, nb, frame, add, c->ads));
    HIPCHK(launch_sum_fixed(x->stream, c->cterm.p, nt, c->vsum.p));
    HIPCHK(hipMemcpyAsync(chi2, c->vsum.p, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    c->fw_seeded[frame] = 1;
    return KS_OK;
}

// _frame_seed! (unbinned): χ² of one frame's image and the dual sweep's seed of that frame on the device
int ks_image_chi2_seed(ks_cache* c, int64_t frame, const double* data, const double* sigma, int sigma_per_pixel, const uint8_t* mask,
                       double* chi2) {
    if (!c || !data || !sigma || !chi2) return fail(KS_ERR_ARG, "ks_image_chi2_seed: null argument");
    if (int rc = seed_frame_ok(c, frame)) return rc;
    const int64_t nsig = sigma_per_pixel ? 4 * c->npix : 4;
    for (int64_t k = 0; k < nsig; ++k)
        if (!(sigma[k] > 0.0) && (!mask || !sigma_per_pixel || mask[k / 4])) return fail(KS_ERR_ARG, "sigma must be positive (entry %lld)", (long long)k);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    if (c->idata.grow((size_t)4 * c->npix) || c->isig.grow((size_t)nsig) || c->vchi.grow(c->npix) || c->vsum.grow(1)) return KS_ERR_HIP;
    if (mask && c->imask.grow(c->npix)) return KS_ERR_HIP;
    HIPCHK(hipMemcpyAsync(c->idata.p, data, (size_t)4 * c->npix * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->isig.p, sigma, (size_t)nsig * sizeof(double), hipMemcpyHostToDevice, x->stream));
    if (mask) HIPCHK(hipMemcpyAsync(c->imask.p, mask, (size_t)c->npix, hipMemcpyHostToDevice, x->stream));
    HIPCHK(launch_image_residual(x->stream, c->out, c->perm, c->npix, frame, c->idata.p, c->isig.p, sigma_per_pixel, mask ? c->imask.p : nullptr,
                                 c->ads, c->vchi.p));
    HIPCHK(launch_sum_fixed(x->stream, c->vchi.p, c->npix, c->vsum.p));
    HIPCHK(hipMemcpyAsync(chi2, c->vsum.p, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    c->fw_seeded[frame] = 1;
    return KS_OK;
}

constexpr bool KS_TANGENT_KERNEL_ENABLED = false;

int ks_spacetime_tangent(ks_cache* c, const ks_model* m, const double* times, int nt, const double* nus, int nnu, double L, int nmax, double slab,
                         double* out_stokes, double* out_tangent) {
    if (!c || !m || !times || !nus || !out_stokes || !out_tangent) return fail(KS_ERR_ARG, "ks_spacetime_tangent: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (m->kind == KS_MODEL_THIN || m->kind == KS_MODEL_KNOT) return fail(KS_ERR_ARG, "the spacetime tangent takes thermal, power-law, kappa or composite models");
    if (nt < 1 || nnu < 1) return fail(KS_ERR_ARG, "need at least one time and one frequency");
    if (!(L > 0.0)) return fail(KS_ERR_ARG, "length unit L must be positive");
    for (int i = 0; i < nnu; ++i) if (!(nus[i] > 0.0)) return fail(KS_ERR_ARG, "frequencies must be positive");
    // The dual kernel (ks_k_tangent.hip) took a GPU memory fault on its first launch and the cause is not
    // found (its item function runs clean on the host under address and undefined-behaviour sanitizers), so
    // the entry point refuses to launch it.
    if (!KS_TANGENT_KERNEL_ENABLED) return fail(KS_ERR_STATE, "ks_spacetime_tangent: the device tangent kernel is disabled");
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const int64_t items = c->npix * nt * nnu;
    std::vector<double> tn(times, times + nt);
    tn.insert(tn.end(), nus, nus + nnu);
    Buf<double> dtn, dout, dtan;
    if (dtn.grow(tn.size()) || dout.grow((size_t)4 * items) || dtan.grow((size_t)12 * items)) { dtn.release(); dout.release(); dtan.release(); return KS_ERR_HIP; }
    int rc = KS_OK;
    hipError_t e = hipMemcpyAsync(dtn.p, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice, x->stream);
    if (e == hipSuccess)
        e = launch_tangent(x->stream, m->kind, c->spin, c->thetao, std::log(L), c->al, c->be, c->npix, c->N, m->d_splats, m->ns, dtn.p, nt,
                           dtn.p + nt, nnu, nmax, slab, dout.p, dtan.p);
    if (e == hipSuccess) e = hipMemcpyAsync(out_stokes, dout.p, (size_t)4 * items * sizeof(double), hipMemcpyDeviceToHost, x->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out_tangent, dtan.p, (size_t)12 * items * sizeof(double), hipMemcpyDeviceToHost, x->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
    if (e != hipSuccess) rc = fail(KS_ERR_HIP, "ks_spacetime_tangent: %s", hipGetErrorString(e));
    dtn.release(); dout.release(); dtan.release();
    return rc;
}

int ks_binning_create(ks_ctx* ctx, int64_t npoints, const int64_t* pixel, int n_alpha, int n_beta, ks_binning** out) {
    if (!ctx || !out || (npoints > 0 && !pixel)) return fail(KS_ERR_ARG, "ks_binning_create: null argument");
    if (n_alpha < 1 || n_beta < 1) return fail(KS_ERR_ARG, "binned screen %d x %d", n_alpha, n_beta);
    const int64_t nbins = (int64_t)n_alpha * n_beta;
    if (npoints > (int64_t)1 << 24 || nbins > npoints) return fail(KS_ERR_DIM, "%lld points for %lld pixels", (long long)npoints, (long long)nbins);
    std::vector<int> off(nbins + 1, 0), pts(npoints), pix(npoints);
    for (int64_t m = 0; m < npoints; ++m) {
        if (pixel[m] < 0 || pixel[m] >= nbins) return fail(KS_ERR_ARG, "pixel index %lld outside 0:%lld", (long long)pixel[m], (long long)(nbins - 1));
        off[pixel[m] + 1] += 1;
        pix[m] = (int)pixel[m];
    }
    for (int64_t q = 0; q < nbins; ++q) {
        if (off[q + 1] == 0) return fail(KS_ERR_ARG, "every pixel needs at least one point (pixel %lld has none)", (long long)q);
        off[q + 1] += off[q];
    }
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int64_t m = 0; m < npoints; ++m) pts[fill[pix[m]]++] = (int)m;       // ascending m within each pixel
    HIPCHK(hipSetDevice(ctx->device));
    ks_binning* b = new ks_binning;
    b->ctx = ctx; b->npoints = npoints; b->nbins = nbins; b->na = n_alpha; b->nb = n_beta;
    if (dalloc(&b->off, off.size()) || dalloc(&b->pts, pts.size()) || dalloc(&b->pix, pix.size())) { ks_binning_destroy(b); return KS_ERR_HIP; }
    HIPCHK(hipMemcpy(b->off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->pts, pts.data(), pts.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->pix, pix.data(), pix.size() * sizeof(int), hipMemcpyHostToDevice));
    *out = b;
    return KS_OK;
}

int ks_binning_destroy(ks_binning* b) {
    if (!b) return KS_OK;
    (void)hipSetDevice(b->ctx->device);
    dfree(b->off); dfree(b->pts); dfree(b->pix);
    delete b;
    return KS_OK;
}

int ks_image_chi2_seed_binned(ks_cache* c, int64_t frame, const ks_binning* b, const double* data, const double* sigma, int sigma_per_pixel,
                              const uint8_t* mask, double* chi2) {
    if (!c || !b || !data || !sigma || !chi2) return fail(KS_ERR_ARG, "ks_image_chi2_seed_binned: null argument");
    if (b->ctx != c->ctx) return fail(KS_ERR_ARG, "the binning belongs to another context");
    if (b->npoints != c->npix) return fail(KS_ERR_DIM, "binning has %lld points, cache has %lld pixels", (long long)b->npoints, (long long)c->npix);
    if (int rc = seed_frame_ok(c, frame)) return rc;
    const int64_t nq = b->nbins;
    const int64_t nsig = sigma_per_pixel ? 4 * nq : 4;
    for (int64_t k = 0; k < nsig; ++k)
        if (!(sigma[k] > 0.0) && (!mask || !sigma_per_pixel || mask[k / 4])) return fail(KS_ERR_ARG, "sigma must be positive (entry %lld)", (long long)k);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    if (c->idata.grow((size_t)4 * nq) || c->isig.grow((size_t)nsig) || c->vchi.grow(nq) || c->vsum.grow(1) || c->bseed.grow((size_t)4 * nq))
        return KS_ERR_HIP;
    if (mask && c->imask.grow(nq)) return KS_ERR_HIP;
    HIPCHK(hipMemcpyAsync(c->idata.p, data, (size_t)4 * nq * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(c->isig.p, sigma, (size_t)nsig * sizeof(double), hipMemcpyHostToDevice, x->stream));
    if (mask) HIPCHK(hipMemcpyAsync(c->imask.p, mask, (size_t)nq, hipMemcpyHostToDevice, x->stream));
    HIPCHK(launch_bin_residual(x->stream, c->out, c->iperm, c->npix, frame, b->off, b->pts, nq, c->idata.p, c->isig.p, sigma_per_pixel,
                               mask ? c->imask.p : nullptr, c->bseed.p, c->vchi.p));
    HIPCHK(launch_bin_seed(x->stream, c->perm, b->pix, c->npix, frame, c->bseed.p, c->ads));
    HIPCHK(launch_sum_fixed(x->stream, c->vchi.p, nq, c->vsum.p));
    HIPCHK(hipMemcpyAsync(chi2, c->vsum.p, sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    c->fw_seeded[frame] = 1;
    return KS_OK;
}

// ---- captured fit iterations (hipGraph) ----------------------------------------------------------------------
// One fit iteration of a single-population model as one graph: the upload of the per-splat constants (made on the
// host from the parameters into pinned staging), the transfer with the records over every (frame, frequency)
// slice, the images, each slice's χ² and seed (image or visibility loss, data resident in graph-owned buffers), the
// reverse pass, and the copies of the gradient rows and per-slice χ² into pinned host memory. A step replays it;
// it is captured again when the model (its splat count), the cache's buffers or its launch choices have changed.
struct ks_graph {
    ks_cache* c = nullptr;
    int nt = 0, nnu = 0, frames = 0;
    double L = 0.0;
    int loss = 0;                                      // KS_LOSS_IMAGE or KS_LOSS_VIS
    Buf<double> dtn, data, sig, chif;
    Buf<uint8_t> mask;
    int sigma_per_pixel = 0;
    // visibility loss: (u, v, taper), phase tables and per-slice workspace
    int nuv = 0; double psize = 0.0; bool tap = false;
    Buf<double> u, v, t, A2, B2, Im, C2, V, wq, E2, Sd, Sp, chi;
    // pinned staging
    SplatC* h_sp = nullptr; SplatAdj* h_ad = nullptr; SplatCf* h_spf = nullptr; SplatAdjf* h_adf = nullptr;
    double* h_grad = nullptr; double* h_chi = nullptr;
    int cap_ns = 0;
    // what the captured graph refers to
    const ks_model* m = nullptr; uint64_t m_id = 0;
    std::vector<const void*> ptrs;
    int gen = -1, prec = -1, chunk = -1, queue = -1, pa = -1, pb = -1;
    hipGraph_t graph = nullptr; hipGraphExec_t exec = nullptr;
    int64_t replays = 0, captures = 0;
};

static void graph_release_capture(ks_graph* g) {
    if (g->exec) { (void)hipGraphExecDestroy(g->exec); g->exec = nullptr; }
    if (g->graph) { (void)hipGraphDestroy(g->graph); g->graph = nullptr; }
}
static void graph_release_host(ks_graph* g) {
    if (g->h_sp) (void)hipHostFree(g->h_sp);
    if (g->h_ad) (void)hipHostFree(g->h_ad);
    if (g->h_spf) (void)hipHostFree(g->h_spf);
    if (g->h_adf) (void)hipHostFree(g->h_adf);
    if (g->h_grad) (void)hipHostFree(g->h_grad);
    g->h_sp = nullptr; g->h_ad = nullptr; g->h_spf = nullptr; g->h_adf = nullptr; g->h_grad = nullptr;
    g->cap_ns = 0;
}

// the cache buffers a captured iteration reads or writes
static std::vector<const void*> graph_ptrs(const ks_cache* c) {
    return {c->seg, c->out, c->arec, c->abar, c->aflg, c->atail, c->awT, c->ads, c->apart, c->adpar, c->aub, c->tmap, c->soa,
            c->qhead.p, c->qorder.p, c->kstop, c->aheads.p, c->sp4};
}

// the iteration's launches on the (capturing) stream
static int graph_record(ks_graph* g, const ks_model* m) {
    ks_cache* c = g->c;
    ks_ctx* x = c->ctx;
    hipStream_t s = x->stream;
    const int ns = m->ns;
    HIPCHK(hipMemcpyAsync(m->d_splats, g->h_sp, ns * sizeof(SplatC), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(m->d_adj, g->h_ad, ns * sizeof(SplatAdj), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(m->d_splats_f, g->h_spf, ns * sizeof(SplatCf), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(m->d_adj_f, g->h_adf, ns * sizeof(SplatAdjf), hipMemcpyHostToDevice, s));
    const int nt = g->nt, nnu = g->nnu;
    if (int rc = launch_transfer(c, m, g->dtn.p, nt, nnu, g->L, -1, 0.0, c->arec, c->aflg)) return rc;
    HIPCHK(launch_seg_combine(s, c->soa, c->npix, c->N, c->M, c->chunk, nt, g->dtn.p + nt, nnu, c->seg, c->out));
    const int na = c->nalpha, nb = c->nbeta;
    for (int f = 0; f < g->frames; ++f) {
        if (g->loss == KS_LOSS_IMAGE) {
            const double* sg = g->sig.p + (g->sigma_per_pixel ? (int64_t)4 * c->npix * f : (int64_t)4 * f);
            HIPCHK(launch_image_residual(s, c->out, c->perm, c->npix, f, g->data.p + (int64_t)4 * c->npix * f, sg, g->sigma_per_pixel,
                                         g->mask.p ? g->mask.p + (int64_t)c->npix * f : nullptr, c->ads, g->chi.p));
            HIPCHK(launch_sum_fixed(s, g->chi.p, c->npix, g->chif.p + f));
        } else {
            const VisWs w{g->u.p, g->v.p, g->t.p, g->A2.p, g->B2.p, g->Im.p, g->C2.p, g->V.p};
            const double scale = g->psize * g->psize / JY;
            const double* d_tap = g->tap ? g->t.p : nullptr;
            if (int rc = vis_forward_ws(c, w, f, 1, g->nuv, d_tap, scale)) return rc;
            HIPCHK(launch_vis_residual(s, g->nuv, g->V.p, g->data.p + (int64_t)8 * g->nuv * f, g->sig.p + (int64_t)4 * g->nuv * f, nullptr, d_tap,
                                       scale, g->wq.p, g->chi.p));
            HIPCHK(launch_vis_rotate(s, g->nuv, nb, g->wq.p, g->B2.p, g->E2.p));
            HIPCHK(launch_dgemm_tn_splitk(s, na, 4 * nb, 2 * g->nuv, g->A2.p, na, g->E2.p, 4 * nb, g->Sd.p, VIS_SPLITK, g->Sp.p));
            HIPCHK(launch_vis_scatter_seed(s, g->Sd.p, c->perm, c->npix, na, nb, f, 0, c->ads));
            HIPCHK(launch_sum_fixed(s, g->chi.p, g->nuv, g->chif.p + f));
        }
    }
    if (int rc = reverse_launches(c, m, g->dtn.p, nt, nnu)) return rc;
    if (ns > 0) HIPCHK(hipMemcpyAsync(g->h_grad, c->adpar, (size_t)ns * ADJ_ROWS * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(g->h_chi, g->chif.p, (size_t)g->frames * sizeof(double), hipMemcpyDeviceToHost, s));
    return KS_OK;
}

// an eager iteration (every buffer allocated at its size), then the capture
static int graph_capture(ks_graph* g, const ks_model* m) {
    ks_cache* c = g->c;
    ks_ctx* x = c->ctx;
    graph_release_capture(g);
    if (g->cap_ns < m->ns || !g->h_sp) {
        graph_release_host(g);
        const size_t n = (size_t)std::max(m->ns, 1);
        HIPCHK(hipHostMalloc((void**)&g->h_sp, n * sizeof(SplatC), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&g->h_ad, n * sizeof(SplatAdj), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&g->h_spf, n * sizeof(SplatCf), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&g->h_adf, n * sizeof(SplatAdjf), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&g->h_grad, n * ADJ_ROWS * sizeof(double), hipHostMallocDefault));
        g->cap_ns = (int)n;
    }
    const int64_t frames = g->frames, items = c->npix * frames;
    if (ensure_out(c, frames) || ensure_adj(c, items, std::max(m->ns, 1))) return KS_ERR_HIP;
    // the eager iteration: the constants already on the device, the same launches, a synchronization
    HIPCHK(hipMemcpy(g->h_sp, m->d_splats, m->ns * sizeof(SplatC), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(g->h_ad, m->d_adj, m->ns * sizeof(SplatAdj), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(g->h_spf, m->d_splats_f, m->ns * sizeof(SplatCf), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(g->h_adf, m->d_adj_f, m->ns * sizeof(SplatAdjf), hipMemcpyDeviceToHost));
    c->fw_chunk = c->chunk;
    if (int rc = graph_record(g, m)) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    // the capture
    HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
    x->capturing = true;
    const int rc = graph_record(g, m);
    x->capturing = false;
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(x->stream, &gr);
    if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
    if (e != hipSuccess) return fail(KS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    g->graph = gr;
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->m = m; g->m_id = m->id;
    g->ptrs = graph_ptrs(c);
    g->gen = c->generation; g->prec = c->precision; g->chunk = c->chunk; g->queue = c->queue; g->pa = c->patch_a; g->pb = c->patch_b;
    g->captures += 1;
    return KS_OK;
}

int ks_fit_graph_create(ks_cache* c, int nt, const double* times, int nnu, const double* nus, double L, int loss, const double* data,
                        const double* sigma, int sigma_per_pixel, const uint8_t* mask, double psize, const double* u, const double* v,
                        const double* taper, int nuv, ks_graph** out) {
    if (!c || !times || !nus || !data || !sigma || !out) return fail(KS_ERR_ARG, "ks_fit_graph_create: null argument");
    if (c->generation < 1) return fail(KS_ERR_STATE, "regenerate! the cache first");
    if (!has_samples(c)) return fail(KS_ERR_ARG, "the cache holds no stored samples: build it with store_samples = true and a storing marcher");
    if (nt < 1 || nnu < 1 || !(L > 0.0)) return fail(KS_ERR_ARG, "need times, frequencies and a positive length unit");
    for (int i = 0; i < nnu; ++i) if (!(nus[i] > 0.0)) return fail(KS_ERR_ARG, "frequencies must be positive");
    if (loss != KS_LOSS_IMAGE && loss != KS_LOSS_VIS) return fail(KS_ERR_ARG, "unknown loss %d", loss);
    if (loss == KS_LOSS_VIS && (!u || !v || nuv < 1)) return fail(KS_ERR_ARG, "the visibility loss needs (u, v) points");
    if (int rc = screen_ready(c)) return rc;
    const int frames = nt * nnu;
    const int64_t nsig = loss == KS_LOSS_VIS ? (int64_t)4 * nuv * frames : sigma_per_pixel ? (int64_t)4 * c->npix * frames : (int64_t)4 * frames;
    for (int64_t k = 0; k < nsig; ++k) {
        const bool masked = loss == KS_LOSS_IMAGE && mask && sigma_per_pixel && !mask[k / 4];
        if (!(sigma[k] > 0.0) && !masked) return fail(KS_ERR_ARG, "sigma must be positive (entry %lld)", (long long)k);
    }
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    ks_graph* g = new ks_graph;
    g->c = c; g->nt = nt; g->nnu = nnu; g->frames = frames; g->L = L; g->loss = loss; g->sigma_per_pixel = sigma_per_pixel;
    std::vector<double> tn(times, times + nt);
    tn.insert(tn.end(), nus, nus + nnu);
    const int64_t ndata = loss == KS_LOSS_VIS ? (int64_t)8 * nuv * frames : (int64_t)4 * c->npix * frames;
    int rc = 0;
    rc |= g->dtn.grow(tn.size()) | g->data.grow(ndata) | g->sig.grow(nsig) | g->chif.grow(frames);
    if (hipHostMalloc((void**)&g->h_chi, frames * sizeof(double), hipHostMallocDefault) != hipSuccess) rc = KS_ERR_HIP;
    if (loss == KS_LOSS_IMAGE) {
        rc |= g->chi.grow(c->npix);
        if (mask) rc |= g->mask.grow((size_t)c->npix * frames);
    } else {
        const int na = c->nalpha, nb = c->nbeta;
        g->nuv = nuv; g->psize = psize; g->tap = taper != nullptr;
        rc |= g->u.grow(nuv) | g->v.grow(nuv) | g->t.grow(nuv) | g->A2.grow((size_t)2 * nuv * na) | g->B2.grow((size_t)2 * nuv * nb) |
              g->Im.grow((size_t)na * 4 * nb) | g->C2.grow((size_t)2 * nuv * 4 * nb) | g->V.grow((size_t)8 * nuv) | g->wq.grow((size_t)8 * nuv) |
              g->E2.grow((size_t)2 * nuv * 4 * nb) | g->Sd.grow((size_t)na * 4 * nb) | g->Sp.grow((size_t)VIS_SPLITK * na * 4 * nb) | g->chi.grow(nuv);
    }
    if (rc) { ks_fit_graph_destroy(g); return fail(KS_ERR_HIP, "ks_fit_graph_create: device allocation failed"); }
#define HIPCHK_G(e) do { hipError_t e_ = (e); if (e_ != hipSuccess) { ks_fit_graph_destroy(g); return fail(KS_ERR_HIP, "%s: %s", #e, hipGetErrorString(e_)); } } while (0)
    HIPCHK_G(hipMemcpy(g->dtn.p, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK_G(hipMemcpy(g->data.p, data, ndata * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK_G(hipMemcpy(g->sig.p, sigma, nsig * sizeof(double), hipMemcpyHostToDevice));
    if (loss == KS_LOSS_IMAGE && mask) HIPCHK_G(hipMemcpy(g->mask.p, mask, (size_t)c->npix * frames, hipMemcpyHostToDevice));
    if (loss == KS_LOSS_VIS) {
        const VisWs w{g->u.p, g->v.p, g->t.p, g->A2.p, g->B2.p, g->Im.p, g->C2.p, g->V.p};
        if (int r2 = vis_tables_ws(c, w, psize, u, v, taper, nuv)) { ks_fit_graph_destroy(g); return r2; }
        HIPCHK_G(hipStreamSynchronize(x->stream));
    }
#undef HIPCHK_G
    *out = g;
    return KS_OK;
}

int ks_fit_graph_step(ks_graph* g, const ks_model* m, const double* params, double* chi2, double* dparams) {
    if (!g || !m || !chi2 || !dparams) return fail(KS_ERR_ARG, "ks_fit_graph_step: null argument");
    ks_cache* c = g->c;
    if (m->ctx != c->ctx) return fail(KS_ERR_ARG, "the model belongs to another context");
    if (m->kind != KS_MODEL_THERMAL && m->kind != KS_MODEL_POWERLAW && m->kind != KS_MODEL_KAPPA && m->kind != KS_MODEL_KNOT)
        return fail(KS_ERR_ARG, "a fit graph takes one thermal, power-law, kappa or knot population (model kind %d)", m->kind);
    ks_ctx* x = c->ctx;
    HIPCHK(hipSetDevice(x->device));
    const bool stale = !g->exec || m != g->m || m->id != g->m_id || graph_ptrs(c) != g->ptrs || c->generation != g->gen ||
                       c->precision != g->prec || c->chunk != g->chunk || c->queue != g->queue || c->patch_a != g->pa || c->patch_b != g->pb;
    if (stale)
        if (int rc = graph_capture(g, m)) return rc;
    // the constants of the new parameters into the pinned staging the graph uploads from
    if (params) {
        const bool with_knots = m->nknots >= 1 && !m->h_knots.empty();
        for (int i = 0; i < m->ns; ++i) {
            host_splat_constants(m->kind, m->nrows, i, params + (size_t)m->nrows * i, m->gamma_max, with_knots, with_knots ? m->h_knots.data() : nullptr,
                                 m->nknots, g->h_sp[i], g->h_ad[i]);
            splat_as(g->h_sp[i], g->h_spf[i]);
            adj_as(g->h_ad[i], g->h_adf[i]);
        }
    }
    HIPCHK(hipGraphLaunch(g->exec, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    g->replays += 1;
    // the forward pass the graph ran is the cache's: images and seeds as after ks_gradient_forward + the seeds
    c->fw_model = m->id; c->fw_nt = g->nt; c->fw_nnu = g->nnu; c->fw_generation = c->generation; c->fw_precision = c->precision;
    c->fw_chunk = c->chunk;
    c->fw_seeded.assign((size_t)g->frames, 1);
    double s = 0.0;
    for (int f = 0; f < g->frames; ++f) s += g->h_chi[f];
    *chi2 = s;
    size_t off = 0;
    for (int i = 0; i < m->ns; ++i) {
        for (int r = 0; r < m->nrows; ++r) dparams[off + r] += g->h_grad[(size_t)i * ADJ_ROWS + r];
        off += m->nrows;
    }
    return KS_OK;
}

int ks_fit_graph_info(const ks_graph* g, int64_t* captures, int64_t* replays) {
    if (!g) return fail(KS_ERR_ARG, "null graph");
    if (captures) *captures = g->captures;
    if (replays) *replays = g->replays;
    return KS_OK;
}

int ks_fit_graph_destroy(ks_graph* g) {
    if (!g) return KS_OK;
    (void)hipSetDevice(g->c->ctx->device);
    graph_release_capture(g);
    graph_release_host(g);
    if (g->h_chi) (void)hipHostFree(g->h_chi);
    for (Buf<double>* b : {&g->dtn, &g->data, &g->sig, &g->chif, &g->u, &g->v, &g->t, &g->A2, &g->B2, &g->Im, &g->C2, &g->V, &g->wq, &g->E2,
                           &g->Sd, &g->Sp, &g->chi})
        b->release();
    g->mask.release();
    delete g;
    return KS_OK;
}

}  // extern "C"

