// Transfer kernels: the polarized transport of every (pixel, time, frequency) item as per-segment
// 4x4 transfer maps, from stored samples (a chunk of samples per work item) or fused with the march (one
// anchor segment per work item), their composition along each ray, and thin (Stokes I) images.
// Reference: transport.jl:99-316, polarized.jl:41-198, Splats.jl:58-147, populations.jl:13-104,
// trajectories.jl:80-128, step.jl:342-398.
#include "ks_kernels.hpp"
#include "ks_dsample.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace ksd {

// ---- polarized transport from stored samples (transport.jl:263-290) ----------------------------------------
// Work items (chunk m, item), item = j + npix (it + nt inu) fastest: samples m chunk + 1 .. of the
// item's ray, composed into the chunk's map seg[(20 m + f) nitems + item]. With rec != nullptr the
// adjoint records (RecField) of every step-taking sample are written, and aflg[(k - 1) nitems + item]
// = 1 when the sample takes a step. A task is one workgroup's 128 consecutive slots (slot = item slot +
// nitems m); the item slots run over the cache's wavefront patches (slot_item).
struct TransferArgs {
    double a, thetao;
    const double* soa;
    int64_t npix;
    int N, chunk, nch;
    Samples smp;
    int ns, nknots;
    const double* times; int nt;
    const double* nus; int nnu;
    double L;
    const int* kstop;
    double* seg;
    uint8_t* aflg;
};
KS_DEV int64_t npix_items(const TransferArgs& A) { return A.npix * A.nt * A.nnu; }
template <int KIND, class S, bool CULL>
__device__ __forceinline__ void transfer_task(const TransferArgs& A, const SplatCT<S>* __restrict__ sp, const S* __restrict__ knots,
                                    const S* __restrict__ ktimes, S* __restrict__ rec, int64_t task, uint16_t* s_cand) {
    const int64_t npix = A.npix;
    const int N = A.N;
    const int64_t nitems = npix * A.nt * A.nnu;
    const int64_t tid = task * 128 + threadIdx.x;
    const bool live = tid < nitems * A.nch;     // the tail's lanes take slot 0 and write nothing
    const int64_t slot = live ? tid : 0;
    const Samples& smp = A.smp;
    const int64_t item = slot_item(smp, slot % nitems, npix);     // wavefronts over compact patches of sky
    const int m = (int)(slot / nitems);
    const int64_t j = item % npix;
    const int64_t fr = item / npix;
    const int it = (int)(fr % A.nt), inu = (int)(fr / A.nt);
    const double a = A.a;
    const Spin Sp = make_spin(a);
    const double dtau = A.soa[(int64_t)PF_TAUTOT * npix + j] / (double)(N + 1);
    const TransCtxT<S> c = make_ctx<S>(a, Sp.rp, Sp.rm, sin(A.thetao), A.soa[(int64_t)PF_ALPHA * npix + j], A.soa[(int64_t)PF_BETA * npix + j],
                                       A.soa[(int64_t)PF_ETA * npix + j], A.soa[(int64_t)PF_LAM * npix + j], sp, A.ns, knots, ktimes, A.nknots,
                                       A.times[it], A.nus[inu], A.L);
    const int ks = A.kstop ? A.kstop[j] : N + 1;
    const double hor = Sp.rp * (1.0 + 1e-3);
    const int kbeg = m * A.chunk + 1, kend = kbeg + A.chunk - 1 < N ? kbeg + A.chunk - 1 : N;
    const int64_t st = (int64_t)N * nitems;
    // the per-tile cull: the splats that can reach the wavefront's samples of this chunk, in ascending
    // order, as a candidate list in LDS (the rest add exact zeros to every sample of the tile)
    constexpr bool cull = CULL;     // the launcher picks CULL for CULL_MIN_NS <= ns <= CAND_MAX
    const int wv = (int)(threadIdx.x / 64), lane = (int)(threadIdx.x % 64);
    int ncand = 0;
    if constexpr (cull) {
        double lo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, hi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
        if (live)
            for (int k = kbeg; k <= kend; ++k) {
                const int64_t o = (int64_t)(k - 1) * npix + j;
                if ((smp.fl[o] & 1) && smp.r[o] > hor && k < ks) {
                    const double v[4] = {smp.x[o], smp.y[o], smp.z[o], c.t_obs - smp.t[o]};
                    for (int q = 0; q < 4; ++q) { lo[q] = fmin(lo[q], v[q]); hi[q] = fmax(hi[q], v[q]); }
                }
            }
        double px = 0.0, py = 0.0, pz = 0.0, R = 0.0, tmin = 0.0, tmax = 0.0;
        const bool any = wave_region(lo, hi, px, py, pz, R, tmin, tmax);
        for (int i0 = 0; i0 < A.ns; i0 += 64) {
            const int i = i0 + lane;
            const bool hit = any && i < A.ns && splat_may_reach(sp[i], px, py, pz, R, tmin, tmax, knots, ktimes);
            const uint64_t mk = __ballot(hit);
            if (hit) s_cand[wv * CAND_MAX + ncand + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u))] = (uint16_t)i;
            ncand += __popcll(mk);
        }
    }
    if constexpr (cull) __syncthreads();
    if (!live) return;
    SegMapT<S> sm;
    sm.init();
    for (int k = kbeg; k <= kend; ++k) {
        const int64_t o = (int64_t)(k - 1) * npix + j;
        const uint8_t fl = smp.fl[o];
        bool step = false;
        const double r = smp.r[o];
        if ((fl & 1) && r > hor && k < ks) {
            const double th = smp.th[o];
            S j4[4], a4[4], r3[3];
            GeoT<S> G;
            if (sample_coefficients<KIND, S>(c, smp.t[o], r, th, (fl & 2) != 0, (fl & 4) != 0, S(smp.x[o]), S(smp.y[o]), S(smp.z[o]),
                                             j4, a4, r3, cull ? s_cand + wv * CAND_MAX : nullptr, ncand, rec ? &G : nullptr)) {
                const S D = S(path_weight(A.L, A.nus[inu], a, r, th, dtau));
                map_step(sm, j4, a4, r3, D);
                step = true;
                if (rec) {
                    S* w = rec + (int64_t)(k - 1) * nitems + item;
                    for (int i = 0; i < 4; ++i) { w[(REC_J + i) * st] = j4[i]; w[(REC_A + i) * st] = a4[i]; }
                    for (int i = 0; i < 3; ++i) w[(REC_RHO + i) * st] = r3[i];
                    w[REC_D * st] = D;
                    geo_store(G, w + REC_GEO * st, st);
                }
            }
        }
        if (rec) A.aflg[(int64_t)(k - 1) * nitems + item] = step ? 1 : 0;
    }
    sm.store(A.seg, m, nitems, item);
}

// The launch forms share one kernel and one copy of the task body (so both give the same numbers bit for bit):
//  * one workgroup per task (head == nullptr): workgroup b runs task q = b, or tasks[b] when a task list is given
//    (the autotuner's sample of the work), ntasks = the grid;
//  * the persistent work queue (head != nullptr): a grid sized to the resident workgroups, each taking the next
//    ticket q from the queue head until ntasks are taken; task q is tasks[q] (a cost-descending order from an
//    earlier launch's per-task times, or a sample of the work) or q itself (ntasks < 2^31).
// With cost != nullptr the wall-clock ticks each task took go to cost[task] (zeroed by the launcher). The loop holds
// no lane-dependent branch outside the task: all 64 lanes of wavefront 0 execute the pop (lane 0 adds one, the others
// zero), the first lane's ticket is shared through LDS, and every exit and back edge of the loop is a
// wavefront-uniform (scalar) branch, so all lanes of the workgroup iterate together and each pass through the
// barriers is a pass through the pop.
template <int KIND, class S, bool CULL>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(Waves<S>::transfer)))
ks_chunk_transfer(TransferArgs A, const SplatCT<S>* __restrict__ sp, const S* __restrict__ knots, const S* __restrict__ ktimes,
                  S* __restrict__ rec, const int* __restrict__ tasks, int64_t ntasks, unsigned* __restrict__ head,
                  uint32_t* __restrict__ cost) {
    __shared__ uint16_t s_cand[2][CULL ? CAND_MAX : 1];
    __shared__ unsigned s_q;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    int64_t next = blockIdx.x;
    for (;;) {
        int64_t q;
        if (head) {
            if (wave == 0) {
                const unsigned ticket = atomicAdd(head, threadIdx.x == 0 ? 1u : 0u);
                s_q = __builtin_amdgcn_readfirstlane(ticket);
            }
            __syncthreads();
            q = (int64_t)__builtin_amdgcn_readfirstlane(s_q);
            __syncthreads();             // every wavefront has read s_q before the next pop overwrites it
        } else {
            q = next;                    // one task, then the loop ends
            next = ntasks;
        }
        if (q >= ntasks) break;
        const int64_t task = tasks ? (int64_t)tasks[q] : q;
        const uint64_t t0 = cost ? wall_clock64() : 0;
        transfer_task<KIND, S, CULL>(A, sp, knots, ktimes, rec, task, &s_cand[0][0]);
        if (cost) {
            __syncthreads();
            atomicMax(cost + task, (uint32_t)min<uint64_t>(wall_clock64() - t0, 0xffffffffull));
        }
    }
}

// ---- the fused march + transport (caches without stored samples) --------------------------------------------
template <int KIND>
struct FusedF {
    TransCtx c;
    int kstop;
    SegMap sm;
    KS_DEV void operator()(int k, const Samp& s, double dtau) {
        if (!(s.ok && s.r > c.rp * (1.0 + 1e-3)) || k >= kstop) return;
        double x, y, z;
        quasi_cartesian_ks(c.a, c.rp, c.rm, s.r, s.th, s.ph, x, y, z);
        double j4[4], a4[4], r3[3];
        if (!sample_coefficients<KIND, double>(c, s.t, s.r, s.th, s.nur, s.nuth, x, y, z, j4, a4, r3)) return;
        sm.step(j4, a4, r3, path_weight(c.L, c.nu, c.a, s.r, s.th, dtau));
    }
};
// Work items (anchor segment m, item), item fastest; segment m of pixel j covers samples m L + 1 ..
// min((m + 1) L, N), L = segment_length of the pixel; items past the pixel's last segment exit.
// Output: seg as ks_chunk_transfer; the segment's anchor residuals res[2 (m npix + j) + {0, 1}] from frame 0.
template <int KIND>
__global__ void __launch_bounds__(128) ks_seg_polarized(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                        const SplatC* __restrict__ sp, int ns,
                                                        const double* __restrict__ knots, const double* __restrict__ ktimes, int nknots,
                                                        const double* __restrict__ times, int nt, const double* __restrict__ nus, int nnu,
                                                        double L, const int* __restrict__ kstop, int nseg_max,
                                                        double* __restrict__ seg, double* __restrict__ res) {
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
    double r1 = 0.0, r2 = 0.0;
    if (kbeg <= N) {
        const int kend = kbeg + Lseg - 1 < N ? kbeg + Lseg - 1 : N;
        const Spin S = make_spin(a);
        FusedF<KIND> f;
        f.c = make_ctx<double>(a, S.rp, S.rm, sin(thetao), p.alpha, p.beta, p.eta, p.lam, sp, ns, knots, ktimes, nknots, times[it], nus[inu], L);
        f.kstop = kstop ? kstop[j] : N + 1;
        f.sm.init();
        const double dtau = p.tau_total / (double)(N + 1);
        // a segment that starts past the outer horizon holds no valid sample and no valid anchor
        if (kbeg * dtau <= p.tau_h) march_range(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
        f.sm.store(seg, m, nitems, item);
    }
    if (fr == 0) { res[2 * (m * npix + j)] = r1; res[2 * (m * npix + j) + 1] = r2; }
}

// ---- thin splats: per chunk (stored samples) or anchor segment (fused) sums --------------------------------------
struct ThinF {
    const SplatC* sp; int ns; double a, rp, rm, t_obs, lam, acc;
    KS_DEV void operator()(int, const Samp& s, double dtau) {
        if (!(s.ok && rp * (1.0 + 1e-3) < s.r && s.r < 1e3)) return;
        double x, y, z;
        quasi_cartesian_ks(a, rp, rm, s.r, s.th, s.ph, x, y, z);
        acc += thin_sample(sp, ns, a, lam, t_obs, s.t, s.r, s.th, x, y, z, dtau);
    }
};
// part[m npix + j] = the thin sum over segment m of pixel j (0 past its last segment); res as above
__global__ void __launch_bounds__(128) ks_thin_parts(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                     Samples smp, const SplatC* __restrict__ sp, int ns, double t_obs, int nseg_max,
                                                     double* __restrict__ part, double* __restrict__ res) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= npix * nseg_max) return;
    const int64_t j = tid % npix;
    const int m = (int)(tid / npix);
    const Spin S = make_spin(a);
    double acc = 0.0, r1 = 0.0, r2 = 0.0;
    if (smp.r) {
        const double lam = soa[(int64_t)PF_LAM * npix + j];
        const double dtau = soa[(int64_t)PF_TAUTOT * npix + j] / (double)(N + 1);
        const int kbeg = m * CHUNK + 1, kend = kbeg + CHUNK - 1 < N ? kbeg + CHUNK - 1 : N;
        for (int k = kbeg; k <= kend; ++k) {
            const int64_t o = (int64_t)(k - 1) * npix + j;
            const double r = smp.r[o];
            if (!((smp.fl[o] & 1) && S.rp * (1.0 + 1e-3) < r && r < 1e3)) continue;
            acc += thin_sample(sp, ns, a, lam, t_obs, smp.t[o], r, smp.th[o], smp.x[o], smp.y[o], smp.z[o], dtau);
        }
    } else {
        const PixC p = load_pixel(soa, npix, j);
        const int Lseg = segment_length(M, 1.0 - p.k);
        const int kbeg = m * Lseg + 1;
        if (kbeg <= N && kbeg * (p.tau_total / (double)(N + 1)) <= p.tau_h) {
            const int kend = kbeg + Lseg - 1 < N ? kbeg + Lseg - 1 : N;
            ThinF f{sp, ns, a, S.rp, S.rm, t_obs, p.lam, 0.0};
            march_range(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
            acc = f.acc;
        }
        res[2 * (m * npix + j)] = r1; res[2 * (m * npix + j) + 1] = r2;
    }
    part[m * npix + j] = acc;
}
__global__ void ks_thin_combine(int64_t npix, int nseg_max, const double* __restrict__ part, double* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    double s = 0.0;
    for (int m = 0; m < nseg_max; ++m) s += part[m * npix + j];
    out[j] = s;
}

// ---- composition of the segment maps ---------------------------------------------------------------------------
// The number of segments of pixel j: fixed chunks (chunk > 0) or the pixel's anchor segments.
KS_DEV int item_segments(const double* soa, int64_t npix, int64_t j, int N, int M, int chunk) {
    const int L = chunk > 0 ? chunk : segment_length(M, 1.0 - soa[(int64_t)PF_K * npix + j]);
    return (N + L - 1) / L;
}
// S = sum_m P_0 ... P_{m-1} S_m, the observed Stokes vector nu-hat^3 S (observed_stokes, transport.jl:311-316)
__global__ void ks_seg_combine(const double* __restrict__ soa, int64_t npix, int N, int M, int chunk, int nt, const double* __restrict__ nus,
                               int nnu, const double* __restrict__ seg, double* __restrict__ out) {
    const int64_t nitems = npix * nt * nnu;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const int64_t j = item % npix;
    const int inu = (int)((item / npix) / nt);
    const int nseg = item_segments(soa, npix, j, N, M, chunk);
    double P[16], Sv[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; ++i) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
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
    }
    const double nh = nus[inu] * (1.0 / NU0);
    const double s3 = nh * nh * nh;
    for (int q = 0; q < 4; ++q) out[4 * item + q] = Sv[q] * s3;
}

// ---- launchers ----
__global__ void ks_fill_i32(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
hipError_t launch_fill_i32(hipStream_t s, int* p, int64_t n, int v) {
    hipLaunchKernelGGL(ks_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
template <int KIND, class S, bool CULL>
static hipError_t chunk_transfer_c(hipStream_t s, const TransferArgs& A, const SplatCT<S>* sp, const S* knots, const S* ktimes, S* rec,
                                   const TransferLaunch& tl) {
    const int64_t all = (npix_items(A) * A.nch + 127) / 128;
    const int64_t ntasks = tl.tasks ? tl.ntasks : all;
    if (ntasks <= 0) return hipSuccess;
    if (ntasks >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    if (!tl.queue) {
        hipLaunchKernelGGL((ks_chunk_transfer<KIND, S, CULL>), dim3((unsigned)ntasks), dim3(128), 0, s, A, sp, knots, ktimes, rec, tl.tasks,
                           ntasks, (unsigned*)nullptr, (uint32_t*)nullptr);
        return hipGetLastError();
    }
    if (!tl.head) return hipErrorInvalidValue;
    static int groups = 0;       // per instantiation
    if (!groups) groups = resident_groups(ks_chunk_transfer<KIND, S, CULL>, 128);
    if (!groups) return hipErrorInvalidValue;
    // the head and the per-task costs start at zero (fill kernels: no memset nodes in a captured graph)
    hipLaunchKernelGGL(ks_fill_i32, dim3(1), dim3(64), 0, s, tl.head, (int64_t)1, 0);
    if (tl.cost) hipLaunchKernelGGL(ks_fill_i32, dim3(grid_for(ntasks, 256)), dim3(256), 0, s, (int*)tl.cost, ntasks, 0);
    const int64_t grid = std::min<int64_t>(ntasks, (int64_t)groups * (tl.group_scale > 0 ? tl.group_scale : 1));
    hipLaunchKernelGGL((ks_chunk_transfer<KIND, S, CULL>), dim3((unsigned)grid), dim3(128), 0, s, A, sp, knots, ktimes, rec, tl.tasks,
                       ntasks, (unsigned*)tl.head, tl.cost);
    return hipGetLastError();
}
template <int KIND, class S>
static hipError_t chunk_transfer_k(hipStream_t s, const TransferArgs& A, const SplatCT<S>* sp, const S* knots, const S* ktimes, S* rec,
                                   const TransferLaunch& tl) {
    // the per-tile cull bounds pattern-rotating splats by their arc and knot splats by the sphere holding their knots
    if (A.ns >= CULL_MIN_NS && A.ns <= CAND_MAX) return chunk_transfer_c<KIND, S, true>(s, A, sp, knots, ktimes, rec, tl);
    return chunk_transfer_c<KIND, S, false>(s, A, sp, knots, ktimes, rec, tl);
}
template <class S>
static hipError_t chunk_transfer_s(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                   const SplatCT<S>* sp, int ns, const S* knots, const S* ktimes, int nknots, const double* times, int nt,
                                   const double* nus, int nnu, double L, const int* kstop, double* seg, S* rec, uint8_t* aflg,
                                   const TransferLaunch& tl) {
    if (tl.chunk < 1) return hipErrorInvalidValue;
    const TransferArgs A{a, thetao, soa, npix, N, tl.chunk, (N + tl.chunk - 1) / tl.chunk, smp, ns, nknots, times, nt, nus, nnu, L, kstop, seg, aflg};
    switch (kind) {
        case MODEL_THERMAL: return chunk_transfer_k<MODEL_THERMAL, S>(s, A, sp, knots, ktimes, rec, tl);
        case MODEL_POWERLAW: return chunk_transfer_k<MODEL_POWERLAW, S>(s, A, sp, knots, ktimes, rec, tl);
        case MODEL_KAPPA: return chunk_transfer_k<MODEL_KAPPA, S>(s, A, sp, knots, ktimes, rec, tl);
        case MODEL_COMPOSITE: return chunk_transfer_k<MODEL_COMPOSITE, S>(s, A, sp, knots, ktimes, rec, tl);
        default: return chunk_transfer_k<MODEL_KNOT, S>(s, A, sp, knots, ktimes, rec, tl);
    }
}
hipError_t launch_chunk_transfer(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                 const SplatC* sp, int ns, const double* knots, const double* ktimes, int nknots,
                                 const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                 double* seg, double* rec, uint8_t* aflg, const TransferLaunch& tl) {
    return chunk_transfer_s<double>(s, kind, a, thetao, soa, npix, N, smp, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, rec,
                                    aflg, tl);
}
hipError_t launch_chunk_transfer_f32(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                     const SplatCf* sp, int ns, const f32* knots, const f32* ktimes, int nknots,
                                     const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                     double* seg, f32* rec, uint8_t* aflg, const TransferLaunch& tl) {
    return chunk_transfer_s<f32>(s, kind, a, thetao, soa, npix, N, smp, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, rec,
                                 aflg, tl);
}
__global__ void ks_iota_i32(int* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}
// The tasks in descending order of their recorded costs (ticks): order[0 ..n) from cost[0 ..n); keys and ids are
// n-long work arrays. Called with tmp = nullptr it sets tmp_bytes to the workspace the sort needs.
hipError_t launch_order_tasks(hipStream_t s, const uint32_t* cost, int64_t n, uint32_t* keys, int* ids, int* order, void* tmp, size_t& tmp_bytes) {
    if (tmp) hipLaunchKernelGGL(ks_iota_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, ids, n);
    return rocprim::radix_sort_pairs_desc(tmp, tmp_bytes, cost, keys, ids, order, (unsigned)n, 0, 32, s);
}
int64_t transfer_tasks(int64_t npix, int nt, int nnu, int N, int chunk) {
    return (npix * nt * nnu * ((N + chunk - 1) / chunk) + 127) / 128;
}
template <int KIND>
static void seg_polarized_k(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, const SplatC* sp, int ns,
                            const double* knots, const double* ktimes, int nknots, const double* times, int nt, const double* nus, int nnu,
                            double L, const int* kstop, double* seg, double* res) {
    const int nseg = max_segments(N);
    const int64_t items = npix * nt * nnu;
    hipLaunchKernelGGL(ks_seg_polarized<KIND>, dim3(grid_for(items * nseg, 128)), dim3(128), 0, s, a, thetao, soa, npix, N, M, sp, ns,
                       knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, nseg, seg, res);
}
hipError_t launch_seg_polarized(hipStream_t s, int kind, double a, double thetao, const double* soa, int64_t npix, int N, int M,
                                const SplatC* sp, int ns, const double* knots, const double* ktimes, int nknots,
                                const double* times, int nt, const double* nus, int nnu, double L, const int* kstop,
                                double* seg, double* res) {
    switch (kind) {
        case MODEL_THERMAL: seg_polarized_k<MODEL_THERMAL>(s, a, thetao, soa, npix, N, M, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, res); break;
        case MODEL_POWERLAW: seg_polarized_k<MODEL_POWERLAW>(s, a, thetao, soa, npix, N, M, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, res); break;
        case MODEL_KAPPA: seg_polarized_k<MODEL_KAPPA>(s, a, thetao, soa, npix, N, M, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, res); break;
        case MODEL_COMPOSITE: seg_polarized_k<MODEL_COMPOSITE>(s, a, thetao, soa, npix, N, M, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, res); break;
        default: seg_polarized_k<MODEL_KNOT>(s, a, thetao, soa, npix, N, M, sp, ns, knots, ktimes, nknots, times, nt, nus, nnu, L, kstop, seg, res); break;
    }
    return hipGetLastError();
}
hipError_t launch_thin_parts(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, Samples smp,
                             const SplatC* sp, int ns, double t_obs, int nseg_max, double* part, double* res) {
    hipLaunchKernelGGL(ks_thin_parts, dim3(grid_for(npix * nseg_max, 128)), dim3(128), 0, s, a, thetao, soa, npix, N, M, smp, sp, ns, t_obs,
                       nseg_max, part, res);
    return hipGetLastError();
}
hipError_t launch_thin_combine(hipStream_t s, int64_t npix, int nseg_max, const double* part, double* out) {
    hipLaunchKernelGGL(ks_thin_combine, dim3(grid_for(npix, 256)), dim3(256), 0, s, npix, nseg_max, part, out);
    return hipGetLastError();
}
hipError_t launch_seg_combine(hipStream_t s, const double* soa, int64_t npix, int N, int M, int chunk, int nt, const double* nus, int nnu,
                              const double* seg, double* out) {
    const int64_t items = npix * nt * nnu;
    hipLaunchKernelGGL(ks_seg_combine, dim3(grid_for(items, 128)), dim3(128), 0, s, soa, npix, N, M, chunk, nt, nus, nnu, seg, out);
    return hipGetLastError();
}

// ---- per-ray splat lists (ray_lists, dual_gradient.jl:283-359) ----------------------------------------------
// One thread per (ray j in sorted order, frame f): hits[i + ns (j + npix f)] = 1 when a sample of ray j before its
// half-orbit cutoff kstop[j] (none when kstop is null) that takes a step (ok, outside the horizon guard) lies inside
// splat i's support sphere at the frame's time (outside_support), else 0; cnt[j + npix f] = the number of hits.
// Splats outer: a splat's test stops at the ray's first sample inside its sphere, and each byte is written once.
__global__ void __launch_bounds__(64) ks_ray_hits(double rp, int64_t npix, int N, Samples smp, const SplatC* __restrict__ sp, int ns,
                                                  const double* __restrict__ times, int nt, const int* __restrict__ kstop,
                                                  uint8_t* __restrict__ hits, int* __restrict__ cnt) {
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= npix * nt) return;
    const int64_t j = item % npix;
    const double t_obs = times[item / npix];
    const int kend = kstop ? (kstop[j] - 1 < N ? kstop[j] - 1 : N) : N;
    const double hor = rp * (1.0 + 1e-3);
    int n = 0;
    for (int i = 0; i < ns; ++i) {
        const SplatC& c = sp[i];
        uint8_t hit = 0;
        for (int k = 1; k <= kend; ++k) {
            const int64_t o = (int64_t)(k - 1) * npix + j;
            if (!(smp.fl[o] & 1) || !(smp.r[o] > hor)) continue;
            double sph, cph;
            if (!outside_support(c, t_obs - smp.t[o], smp.x[o], smp.y[o], smp.z[o], sph, cph)) { hit = 1; break; }
        }
        hits[(int64_t)i + (int64_t)ns * item] = hit;
        n += hit;
    }
    cnt[item] = n;
}
// ids[s + cap t] = the (s+1)-th hit splat of output item t in ascending order (0-based), -1 past its count. Output
// item t = p + npix f reads item iperm[p] + npix f of hits (the screen pixel p's sorted slot), or t itself when
// iperm is null.
__global__ void ks_ray_fill(int64_t items, int64_t npix, int ns, int cap, const uint8_t* __restrict__ hits, const int* __restrict__ iperm,
                            int32_t* __restrict__ ids) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    const int64_t it = iperm ? iperm[t % npix] + npix * (t / npix) : t;
    int n = 0;
    for (int i = 0; i < ns; ++i)
        if (hits[(int64_t)i + (int64_t)ns * it] && n < cap) ids[(int64_t)cap * t + n++] = i;
    for (; n < cap; ++n) ids[(int64_t)cap * t + n] = -1;
}

hipError_t launch_ray_hits(hipStream_t s, double rp, int64_t npix, int N, Samples smp, const SplatC* sp, int ns, const double* times, int nt,
                           const int* kstop, uint8_t* hits, int* cnt) {
    hipLaunchKernelGGL(ks_ray_hits, dim3(grid_for(npix * nt, 64)), dim3(64), 0, s, rp, npix, N, smp, sp, ns, times, nt, kstop, hits, cnt);
    return hipGetLastError();
}
hipError_t launch_ray_fill(hipStream_t s, int64_t items, int64_t npix, int ns, int cap, const uint8_t* hits, const int* iperm, int32_t* ids) {
    hipLaunchKernelGGL(ks_ray_fill, dim3(grid_for(items, 128)), dim3(128), 0, s, items, npix, ns, cap, hits, iperm, ids);
    return hipGetLastError();
}

}  // namespace ksd
