// Adjoint kernels: the dual sweep over stored samples (polarized_tails!, polarized_dual_sweep!,
// dual_gradient.jl:125-249): tails and adjoints at the chunk boundaries from the chunk maps, the
// per-sample adjoints inside each chunk, the per-splat element adjoints with a wavefront sum, and a
// deterministic reduction over the item blocks.
#include "ks_kernels.hpp"

namespace ksd {

// ---- the dual-sweep adjoint (dual_gradient.jl:125-249) over stored samples -------------------------------------
// Per item: from the chunk maps (P_m, S_m), the tail behind each chunk T_out[m] = S_{m+1} + P_{m+1} T_out[m+1]
// (T_out[last] = 0), the image nu-hat^3 T_in[0], and the adjoint in front of each chunk
// w_in[m] = (P_0 ... P_{m-1})^T w_1 with w_1 = nu-hat^3 dstokes (dual_gradient.jl:162).
// wT[(8 m + q) nitems + item] = w_in, wT[(8 m + 4 + q) nitems + item] = T_out; the image only when out is given.
__global__ void ks_adj_combine(int64_t npix, int N, int chunk, int nt, const double* __restrict__ nus, int nnu, const double* __restrict__ seg,
                               const double* __restrict__ dstokes, double* __restrict__ wT, double* __restrict__ out) {
    const int64_t nitems = npix * nt * nnu;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    const int inu = (int)((item / npix) / nt);
    const int nseg = (N + chunk - 1) / chunk;
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
    if (out)
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

// The FP64 step and step adjoint of an FP32 sample, out of line (the rare samples with a Faraday depth
// |rho| Delta above FARADAY_FP64, where the forward pass takes the FP64 step too).
// (inputs and outputs by value, as step_f64_for_f32's: the callers' arrays stay in registers)
struct Vec4_32 { f32 v[4]; };
struct Seeds32 { f32 w[4], R[4]; };
struct StepAdj32 { f32 Otw[4], jb[4], ab[4], rb[3]; };
KS_COLD StepAdj32 step_adjoint_f64_for_f32(const Step32 s, const Seeds32 z) {
    double jd[4], ad[4], rd[3], wd[4], Rd[4], Od[4], jbd[4], abd[4], rbd[3];
    for (int i = 0; i < 4; ++i) { jd[i] = val(s.j[i]); ad[i] = val(s.a[i]); wd[i] = val(z.w[i]); Rd[i] = val(z.R[i]); }
    for (int i = 0; i < 3; ++i) rd[i] = val(s.r[i]);
    sample_adjoint(jd, ad, rd, val(s.D), wd, Rd, Od, jbd, abd, rbd);
    StepAdj32 o;
    for (int i = 0; i < 4; ++i) { o.Otw[i] = f32(Od[i]); o.jb[i] = f32(jbd[i]); o.ab[i] = f32(abd[i]); }
    for (int i = 0; i < 3; ++i) o.rb[i] = f32(rbd[i]);
    return o;
}
KS_COLD Vec4_32 step_apply_f64_for_f32(const Step32 s, const Vec4_32 Tin) {
    double jd[4], ad[4], rd[3], Ti[4], To[4];
    for (int i = 0; i < 4; ++i) { jd[i] = val(s.j[i]); ad[i] = val(s.a[i]); Ti[i] = val(Tin.v[i]); }
    for (int i = 0; i < 3; ++i) rd[i] = val(s.r[i]);
    step_apply(jd, ad, rd, val(s.D), Ti, To);
    Vec4_32 o;
    for (int i = 0; i < 4; ++i) o.v[i] = f32(To[i]);
    return o;
}

// The adjoint of one sample's step, and the step applied to a tail, in the transfer's scalar; in FP32
// the samples with a Faraday depth |rho| Delta above FARADAY_FP64 take the FP64 step and its adjoint.
template <class S>
KS_DEV void step_adjoint_s(const S j4[4], const S a4[4], const S r3[3], S D, const S w[4], const S R[4], S Otw[4], S jb[4], S ab[4], S rb[3]) {
    if constexpr (!std::is_same<S, double>::value) {
        const S fd = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]) * D;
        if (fd > FARADAY_FP64) {
            Seeds32 z;
            for (int i = 0; i < 4; ++i) { z.w[i] = w[i]; z.R[i] = R[i]; }
            const StepAdj32 o = step_adjoint_f64_for_f32(step32_of(j4, a4, r3, D), z);
            for (int i = 0; i < 4; ++i) { Otw[i] = o.Otw[i]; jb[i] = o.jb[i]; ab[i] = o.ab[i]; }
            for (int i = 0; i < 3; ++i) rb[i] = o.rb[i];
            return;
        }
    }
    sample_adjoint(j4, a4, r3, D, w, R, Otw, jb, ab, rb);
}
template <class S>
KS_DEV void step_apply_s(const S j4[4], const S a4[4], const S r3[3], S D, const S Tin[4], S Tout[4]) {
    if constexpr (!std::is_same<S, double>::value) {
        const S fd = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]) * D;
        if (fd > FARADAY_FP64) {
            Vec4_32 ti;
            for (int i = 0; i < 4; ++i) ti.v[i] = Tin[i];
            const Vec4_32 o = step_apply_f64_for_f32(step32_of(j4, a4, r3, D), ti);
            for (int i = 0; i < 4; ++i) Tout[i] = o.v[i];
            return;
        }
    }
    step_apply(j4, a4, r3, D, Tin, Tout);
}

// Per (chunk, item): the tails T_{k+1} of the chunk's samples back to front from T_out[m], then the
// adjoint front to back from w_in[m]: per sample the adjoints of its j, alpha, rho (into bar, beside the
// records, which stay intact so that the reverse pass can run again on the same forward pass) and
// w <- O^T w (polarized_dual_kernel!'s ray loop, dual_gradient.jl:151-178).
template <class S>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(Waves<S>::adjstep))) ks_adj_segment(int N, int chunk, int64_t nitems, int nch, const double* __restrict__ wT, const S* __restrict__ rec,
                                                      S* __restrict__ bar, const uint8_t* __restrict__ aflg, S* __restrict__ tails, Samples smp,
                                                      int64_t npix) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nitems * nch) return;
    const int64_t item = slot_item(smp, tid % nitems, npix);     // wavefronts over compact patches of sky
    const int m = (int)(tid / nitems);
    const int kbeg = m * chunk + 1, kend = kbeg + chunk - 1 < N ? kbeg + chunk - 1 : N;
    const int64_t st = (int64_t)N * nitems;
    S T[4], w[4];
    for (int q = 0; q < 4; ++q) { w[q] = S(wT[(8 * (int64_t)m + q) * nitems + item]); T[q] = S(wT[(8 * (int64_t)m + 4 + q) * nitems + item]); }
    for (int k = kend; k >= kbeg; --k) {
        const int64_t fo = (int64_t)(k - 1) * nitems + item;
        if (!aflg[fo]) continue;
        const S* o = rec + fo;
        S j4[4], a4[4], r3[3];
        for (int i = 0; i < 4; ++i) { j4[i] = o[(REC_J + i) * st]; a4[i] = o[(REC_A + i) * st]; }
        for (int i = 0; i < 3; ++i) r3[i] = o[(REC_RHO + i) * st];
        for (int q = 0; q < 4; ++q) tails[(int64_t)q * st + fo] = T[q];
        S Tn[4];
        step_apply_s(j4, a4, r3, o[REC_D * st], T, Tn);
        for (int q = 0; q < 4; ++q) T[q] = Tn[q];
    }
    for (int k = kbeg; k <= kend; ++k) {
        const int64_t fo = (int64_t)(k - 1) * nitems + item;
        if (!aflg[fo]) continue;
        const S* o = rec + fo;
        S j4[4], a4[4], r3[3], R[4];
        for (int i = 0; i < 4; ++i) { j4[i] = o[(REC_J + i) * st]; a4[i] = o[(REC_A + i) * st]; R[i] = tails[(int64_t)i * st + fo]; }
        for (int i = 0; i < 3; ++i) r3[i] = o[(REC_RHO + i) * st];
        S Otw[4], jb[4], ab[4], rb[3];
        step_adjoint_s(j4, a4, r3, o[REC_D * st], w, R, Otw, jb, ab, rb);
        S* b = bar + fo;
        for (int i = 0; i < 4; ++i) { b[i * st] = jb[i]; b[(4 + i) * st] = ab[i]; }
        for (int i = 0; i < 3; ++i) b[(8 + i) * st] = rb[i];
        for (int q = 0; q < 4; ++q) w[q] = Otw[q];
    }
}

// The region of each element-adjoint unit (64 items x ADJ_SLICE samples): the bounding sphere of its
// step-taking samples and their emission times, ub[UB_FIELDS unit + f] (wave_region); one wavefront per unit.
__global__ void __launch_bounds__(64) ks_unit_bounds(int64_t npix, int N, Samples smp, const double* __restrict__ times, int nt,
                                                     int64_t nitems, const uint8_t* __restrict__ aflg, double* __restrict__ ub) {
    const int64_t unit = blockIdx.x;
    const int nslice = (N + ADJ_SLICE - 1) / ADJ_SLICE;
    const int64_t ib = unit / nslice;
    const int kfirst = (int)(unit % nslice) * ADJ_SLICE + 1;
    const int klast = kfirst + ADJ_SLICE - 1 < N ? kfirst + ADJ_SLICE - 1 : N;
    const int64_t item = ib * 64 + threadIdx.x;
    double lo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, hi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (item < nitems) {
        const int64_t j = item % npix;
        const double t_obs = times[(item / npix) % nt];
        for (int k = kfirst; k <= klast; ++k) {
            if (!aflg[(int64_t)(k - 1) * nitems + item]) continue;
            const int64_t so = (int64_t)(k - 1) * npix + j;
            const double v[4] = {smp.x[so], smp.y[so], smp.z[so], t_obs - smp.t[so]};
            for (int q = 0; q < 4; ++q) { lo[q] = fmin(lo[q], v[q]); hi[q] = fmax(hi[q], v[q]); }
        }
    }
    double px = 0.0, py = 0.0, pz = 0.0, R = 0.0, tmin = 0.0, tmax = 0.0;
    const bool any = wave_region(lo, hi, px, py, pz, R, tmin, tmax);
    if (threadIdx.x == 0) {
        double* o = ub + UB_FIELDS * unit;
        o[0] = px; o[1] = py; o[2] = pz; o[3] = R; o[4] = tmin; o[5] = tmax; o[6] = any ? 1.0 : 0.0; o[7] = 0.0;
    }
}

// Per (64 consecutive items, slice of ADJ_SLICE samples, splat): the element adjoints of every
// (item, sample) pair inside the splat's weight cutoff, 21 partial sums per lane, summed over the
// wavefront; part[(u ns + i) 21 + r] for the unit u = ib nslice + slice. The slices bound the work of
// one wavefront: the pairs of a splat crowd into a few item blocks and stretches of the rays, and
// whole-ray units left most of the chip idle behind the heaviest ones.
// The wavefront first walks its items' samples with the cheap weight test, WALK samples at a time with
// their loads issued together, and queues the pairs that pass (ballot + mbcnt into an LDS ring), then
// runs the element adjoint 64 pairs at a time, so that lanes outside the splat's support do not idle
// through it. Every pair of a wavefront belongs to
// splat i, so the sums stay per lane in registers; the queue order is fixed, so the result is
// deterministic.
// POPS: the populations the launch meets, POP_THERMAL (thermal and knot splats), POP_POWERLAW, POP_KAPPA, and
// POP_KNOTS (centres that follow knots); a launch of one population carries only its branch, whose registers would
// otherwise cost it occupancy; mixed populations take a runtime branch on the splat's kind.
template <class S, int POPS>
__device__ __forceinline__ void adj_splats_block(int64_t blk, double a, double thetao, const double* __restrict__ soa, int64_t npix, int N,
                                                 Samples smp, const SplatCT<S>* __restrict__ sp, const SplatAdjT<S>* __restrict__ sadj,
                                                 int ns, const S* __restrict__ knots, const S* __restrict__ ktimes,
                                                 const double* __restrict__ times, int nt, const double* __restrict__ nus, int64_t nitems,
                                                 const S* __restrict__ rec, const S* __restrict__ bar, const uint8_t* __restrict__ aflg,
                                                 int64_t unit0, int64_t nb, const double* __restrict__ ub, double* __restrict__ part) {
    constexpr int WALK = 8;                 // samples tested per group, their loads issued together
    constexpr int QCAP = 1024;              // ring capacity: < 64 pending + 64 per sample of a group
    static_assert(QCAP >= 64 * (WALK + 1) && (QCAP & (QCAP - 1)) == 0, "the ring must hold a group's pushes");
    __shared__ double s_alpha[64], s_beta[64], s_tobs[64], s_nu[64];
    __shared__ int64_t s_item[64];
    __shared__ int s_q[QCAP];   // entry = lane | k << 6
    // XCD-aware order: the hardware deals consecutive blocks round-robin over the 8 XCDs, so block b runs on
    // XCD b % 8; all ns splats of a unit go to one XCD (and its L2, which then reads the unit's samples once)
    // this launch covers the units unit0 .. unit0 + nb - 1 (a batch: the partial sums of a batch are bounded by
    // nb x ns x ADJ_ROWS doubles); `local` numbers the unit within the batch
    int i;
    int64_t unit, local;
    {
        const int64_t b = blk, q = b / 8, x = b % 8;
        i = (int)(q % ns);
        local = (q / ns) * 8 + x;
        if (local >= nb) return;          // the padding of the last group of 8 units
        unit = unit0 + local;
    }
    // the per-unit cull: a splat that cannot reach any step-taking sample of the unit adds zeros
    {
        const double* u = ub + UB_FIELDS * unit;
        if (!(u[6] > 0.0) || !splat_may_reach(sp[i], u[0], u[1], u[2], u[3], u[4], u[5], knots, ktimes)) {
            // every lane stores (lanes past the rows repeat the last row's zero): no lane-dependent branch
            part[(local * ns + i) * ADJ_ROWS + min((int)threadIdx.x, ADJ_ROWS - 1)] = 0.0;
            return;
        }
    }
    const int nslice = (N + ADJ_SLICE - 1) / ADJ_SLICE;
    const int64_t ib = unit / nslice;
    const int kfirst = (int)(unit % nslice) * ADJ_SLICE + 1;
    const int klast = kfirst + ADJ_SLICE - 1 < N ? kfirst + ADJ_SLICE - 1 : N;
    const int lane = (int)threadIdx.x;
    const bool valid = ib * 64 + lane < nitems;
    const int64_t item = ib * 64 + lane;      // sorted order: the pair queue already packs the lanes
    // the splat's constants in LDS, read where they are used: held in registers for the whole kernel
    // they outgrow the scalar file and spill into vector lanes
    __shared__ SplatCT<S> s_c;
    __shared__ SplatAdjT<S> s_sa;
    if (lane == 0) { s_c = sp[i]; s_sa = sadj[i]; }
    const SplatCT<S>& c = s_c;
    const SplatAdjT<S>& sa = s_sa;
    int64_t j = 0;
    double t_obs = 0.0;
    if (valid) {
        j = item % npix;
        const int64_t fr = item / npix;
        t_obs = times[fr % nt];
        s_alpha[lane] = soa[(int64_t)PF_ALPHA * npix + j];
        s_beta[lane] = soa[(int64_t)PF_BETA * npix + j];
        s_tobs[lane] = t_obs;
        s_nu[lane] = nus[fr / nt];
        s_item[lane] = item;
    }
    __syncthreads();
    const S tos = S(sin(thetao)), aS = S(a);
    const int64_t st = (int64_t)N * nitems;
    // the FP32 path reads the samples' FP32 copy (one 16-byte load per sample instead of four 8-byte loads; its
    // launcher requires it)
    constexpr bool use_p4 = std::is_same<S, f32>::value;
    S grad[ADJ_ROWS];
    for (int r = 0; r < ADJ_ROWS; ++r) grad[r] = S(0.0);
    int qhead = 0, qn = 0;   // wavefront-uniform ring state
    auto run_batch = [&](int cnt) {
        if (lane < cnt) {
            const int e = s_q[(qhead + lane) & (QCAP - 1)];
            const int l = e & 63, k = e >> 6;
            const int64_t it2 = s_item[l];
            const int64_t fo = (int64_t)(k - 1) * nitems + it2;
            const int64_t so = (int64_t)(k - 1) * npix + it2 % npix;
            S dl[13];
            S tt, xx, yy, zz;
            if constexpr (use_p4) {
                const float4 q = smp.p4[so];
                tt = S((float)s_tobs[l] - q.w); xx = S(q.x); yy = S(q.y); zz = S(q.z);
            } else {
                tt = S(s_tobs[l] - smp.t[so]); xx = S(smp.x[so]); yy = S(smp.y[so]); zz = S(smp.z[so]);
            }
            S Gw;
            if constexpr ((POPS & POP_KNOTS) != 0) Gw = c.nk > 0 ? weight_dlog_knot(c, sa, knots, ktimes, tt, xx, yy, zz, dl) : weight_dlog(c, sa, tt, xx, yy, zz, dl);
            else Gw = weight_dlog(c, sa, tt, xx, yy, zz, dl);
            if (Gw > 0.0) {
                const S* o = bar + fo;
                S bars[BAR_FIELDS];
                for (int q = 0; q < BAR_FIELDS; ++q) bars[q] = o[q * st];
                GeoT<S> G;
                geo_load(rec + fo + REC_GEO * st, st, G);        // written by the forward transfer
                constexpr int fl = POPS & POP_FLUIDS;
                if constexpr (fl == POP_THERMAL)
                    element_adjoint_thermal(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
                else if constexpr (fl == POP_POWERLAW)
                    element_adjoint_pop(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
                else if constexpr (fl == POP_KAPPA)
                    element_adjoint_kappa(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
                else if (c.kind == MODEL_KAPPA)
                    element_adjoint_kappa(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
                else if (c.kind == MODEL_POWERLAW)
                    element_adjoint_pop(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
                else
                    element_adjoint_thermal(c, sa, G, aS, S(s_alpha[l]), S(s_beta[l]), tos, S(s_nu[l]), bars, Gw, dl, grad);
            }
        }
        qhead = (qhead + cnt) & (QCAP - 1);
        qn -= cnt;
    };
    for (int k0 = kfirst; k0 <= klast; k0 += WALK) {
        uint8_t fl[WALK];
        S tk[WALK], xk[WALK], yk[WALK], zk[WALK];     // t_obs - t~ and the position of each sample of the group
#pragma unroll
        for (int u = 0; u < WALK; ++u) {
            const int k = k0 + u;
            fl[u] = 0;
            tk[u] = xk[u] = yk[u] = zk[u] = S(0.0);
            if (valid && k <= klast) {
                const int64_t so = (int64_t)(k - 1) * npix + j;
                fl[u] = aflg[(int64_t)(k - 1) * nitems + item];
                if constexpr (use_p4) {
                    const float4 q = smp.p4[so];
                    tk[u] = S((float)t_obs - q.w); xk[u] = S(q.x); yk[u] = S(q.y); zk[u] = S(q.z);
                } else {
                    tk[u] = S(t_obs - smp.t[so]); xk[u] = S(smp.x[so]); yk[u] = S(smp.y[so]); zk[u] = S(smp.z[so]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < WALK; ++u) {
            bool act;
            if constexpr ((POPS & POP_KNOTS) != 0)
                act = fl[u] && (c.nk > 0 ? weight_active_knot(c, knots, ktimes, tk[u], xk[u], yk[u], zk[u])
                                         : weight_active(c, tk[u], xk[u], yk[u], zk[u]));
            else
                act = fl[u] && weight_active(c, tk[u], xk[u], yk[u], zk[u]);
            const uint64_t m = __ballot(act);
            if (act) {
                const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                s_q[(qhead + qn + rank) & (QCAP - 1)] = lane | ((k0 + u) << 6);
            }
            qn += __popcll(m);
        }
        while (qn >= 64) {
            __syncthreads();
            run_batch(64);
            __syncthreads();
        }
    }
    if (qn > 0) {
        __syncthreads();
        run_batch(qn);
    }
    // wavefront sum (blockDim = 64 = one wavefront), in FP64
    double gsum[ADJ_ROWS];
    for (int r = 0; r < ADJ_ROWS; ++r) {
        double v = val(grad[r]);
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        gsum[r] = v;
    }
    // lane r < ADJ_ROWS stores row r (the lanes past the rows repeat the last row): selects, no lane-dependent branch
    double mine = gsum[ADJ_ROWS - 1];
    for (int r = 0; r < ADJ_ROWS - 1; ++r) mine = lane == r ? gsum[r] : mine;
    part[(local * ns + i) * ADJ_ROWS + min(lane, ADJ_ROWS - 1)] = mine;
}

// The launch forms of the element adjoints: one workgroup (one wavefront) per block b of the XCD-aware numbering
// (heads == nullptr), or a persistent grid of the resident workgroups (a multiple of 8, so workgroup w runs on XCD
// w % 8) in which each workgroup takes tickets t from its XCD's queue heads[w % 8] and runs block b = 8 t + w % 8
// until the blocks (nblocks) are taken: the blocks of a unit stay on one XCD as in the one-per-block form, and the
// units whose splat is culled cost a ticket instead of a dispatch. The ticket is taken by all 64 lanes (lane 0 adds
// one) and made scalar, so the loop's exit and back edge are uniform branches.
template <class S, int POPS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(Waves<S>::adjelem)))
ks_adj_splats(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, Samples smp, const SplatCT<S>* __restrict__ sp,
              const SplatAdjT<S>* __restrict__ sadj, int ns, const S* __restrict__ knots, const S* __restrict__ ktimes,
              const double* __restrict__ times, int nt, const double* __restrict__ nus, int64_t nitems, const S* __restrict__ rec,
              const S* __restrict__ bar, const uint8_t* __restrict__ aflg, int64_t unit0, int64_t nb, const double* __restrict__ ub,
              double* __restrict__ part, int64_t nblocks, unsigned* __restrict__ heads) {
    const int x = (int)(blockIdx.x % 8);
    int64_t next = blockIdx.x;
    for (;;) {
        int64_t blk;
        if (heads) {
            const unsigned t = __builtin_amdgcn_readfirstlane(atomicAdd(heads + x, threadIdx.x == 0 ? 1u : 0u));
            blk = 8 * (int64_t)t + x;
        } else {
            blk = next;
            next = nblocks;
        }
        if (blk >= nblocks) break;
        adj_splats_block<S, POPS>(blk, a, thetao, soa, npix, N, smp, sp, sadj, ns, knots, ktimes, times, nt, nus, nitems, rec, bar, aflg,
                                  unit0, nb, ub, part);
    }
}

// dparams[r + ADJ_ROWS i] = the sum over a batch's work units in a fixed order (deterministic): one wavefront per
// (splat, row), lane l summing units l, l + 64, ..., then a fixed butterfly over the lanes; added to the sums of
// the earlier batches when accumulate (the batches in order)
__global__ void __launch_bounds__(64) ks_adj_reduce(const double* __restrict__ part, int64_t nunits, int ns, int accumulate,
                                                    double* __restrict__ dparams) {
    const int64_t t = blockIdx.x;
    if (t >= (int64_t)ns * ADJ_ROWS) return;
    const int i = (int)(t / ADJ_ROWS), r = (int)(t % ADJ_ROWS);
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nunits; b += 64) s += part[(b * ns + i) * ADJ_ROWS + r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) dparams[t] = accumulate ? dparams[t] + s : s;
}

// ---- launchers ----
hipError_t launch_adj_combine(hipStream_t s, int64_t npix, int N, int chunk, int nt, const double* nus, int nnu, const double* seg,
                              const double* dstokes, double* wT, double* out) {
    const int64_t items = npix * nt * nnu;
    hipLaunchKernelGGL(ks_adj_combine, dim3(grid_for(items, 128)), dim3(128), 0, s, npix, N, chunk, nt, nus, nnu, seg, dstokes, wT, out);
    return hipGetLastError();
}
hipError_t launch_adj_segment(hipStream_t s, int N, int chunk, int64_t nitems, const double* wT, const double* rec, double* bar,
                              const uint8_t* aflg, double* tails, Samples smp, int64_t npix) {
    const int nch = (N + chunk - 1) / chunk;
    hipLaunchKernelGGL(ks_adj_segment<double>, dim3(grid_for(nitems * nch, 128)), dim3(128), 0, s, N, chunk, nitems, nch, wT, rec, bar, aflg,
                       tails, smp, npix);
    return hipGetLastError();
}
hipError_t launch_adj_segment_f32(hipStream_t s, int N, int chunk, int64_t nitems, const double* wT, const f32* rec, f32* bar,
                                  const uint8_t* aflg, f32* tails, Samples smp, int64_t npix) {
    const int nch = (N + chunk - 1) / chunk;
    hipLaunchKernelGGL(ks_adj_segment<f32>, dim3(grid_for(nitems * nch, 128)), dim3(128), 0, s, N, chunk, nitems, nch, wT, rec, bar, aflg,
                       tails, smp, npix);
    return hipGetLastError();
}
template <class S>
static hipError_t adj_splats_k(hipStream_t s, int pops, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                               const SplatCT<S>* sp, const SplatAdjT<S>* sadj, int ns, const S* knots, const S* ktimes,
                               const double* times, int nt, const double* nus, int64_t nitems, const S* rec, const S* bar,
                               const uint8_t* aflg, int64_t unit0, int64_t nb, const double* ub, double* part, int* heads) {
    const int64_t nblocks = (nb + 7) / 8 * 8 * ns;
    dim3 grid((unsigned)nblocks), block(64);
    unsigned* hq = nullptr;
    if (heads) {
        // the persistent grid: the resident workgroups of the launch's instantiation, a multiple of 8
        int groups = 0;
#define KS_ADJ_GROUPS(P) groups = resident_groups(ks_adj_splats<S, P>, 64)
        switch (pops) {
            case POP_THERMAL: KS_ADJ_GROUPS(POP_THERMAL); break;
            case POP_THERMAL | POP_KNOTS: KS_ADJ_GROUPS(POP_THERMAL | POP_KNOTS); break;
            case POP_POWERLAW: KS_ADJ_GROUPS(POP_POWERLAW); break;
            case POP_POWERLAW | POP_KNOTS: KS_ADJ_GROUPS(POP_POWERLAW | POP_KNOTS); break;
            case POP_KAPPA: KS_ADJ_GROUPS(POP_KAPPA); break;
            case POP_KAPPA | POP_KNOTS: KS_ADJ_GROUPS(POP_KAPPA | POP_KNOTS); break;
            default: KS_ADJ_GROUPS(POP_FLUIDS | POP_KNOTS); break;
        }
#undef KS_ADJ_GROUPS
        if (groups < 8) return hipErrorInvalidValue;
        grid = dim3((unsigned)std::min<int64_t>(groups / 8 * 8, nblocks));
        if (hipError_t e = launch_fill_i32(s, heads, 8, 0)) return e;
        hq = (unsigned*)heads;
    }
#define KS_ADJ_LAUNCH(P) hipLaunchKernelGGL((ks_adj_splats<S, P>), grid, block, 0, s, a, thetao, soa, npix, N, smp, sp, sadj, ns, knots, ktimes, \
                                            times, nt, nus, nitems, rec, bar, aflg, unit0, nb, ub, part, nblocks, hq)
    switch (pops) {
        case POP_THERMAL: KS_ADJ_LAUNCH(POP_THERMAL); break;
        case POP_THERMAL | POP_KNOTS: KS_ADJ_LAUNCH(POP_THERMAL | POP_KNOTS); break;
        case POP_POWERLAW: KS_ADJ_LAUNCH(POP_POWERLAW); break;
        case POP_POWERLAW | POP_KNOTS: KS_ADJ_LAUNCH(POP_POWERLAW | POP_KNOTS); break;
        case POP_KAPPA: KS_ADJ_LAUNCH(POP_KAPPA); break;
        case POP_KAPPA | POP_KNOTS: KS_ADJ_LAUNCH(POP_KAPPA | POP_KNOTS); break;
        default: KS_ADJ_LAUNCH(POP_FLUIDS | POP_KNOTS); break;        // more than one population
    }
#undef KS_ADJ_LAUNCH
    return hipGetLastError();
}
hipError_t launch_adj_splats(hipStream_t s, int pops, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                             const SplatC* sp, const SplatAdj* sadj, int ns, const double* knots, const double* ktimes,
                             const double* times, int nt, const double* nus, int64_t nitems, const double* rec, const double* bar,
                             const uint8_t* aflg, int64_t unit0, int64_t nb, const double* ub, double* part, int* heads) {
    return adj_splats_k<double>(s, pops, a, thetao, soa, npix, N, smp, sp, sadj, ns, knots, ktimes, times, nt, nus, nitems, rec, bar, aflg, unit0,
                                nb, ub, part, heads);
}
hipError_t launch_adj_splats_f32(hipStream_t s, int pops, double a, double thetao, const double* soa, int64_t npix, int N, Samples smp,
                                 const SplatCf* sp, const SplatAdjf* sadj, int ns, const f32* knots, const f32* ktimes,
                                 const double* times, int nt, const double* nus, int64_t nitems, const f32* rec, const f32* bar,
                                 const uint8_t* aflg, int64_t unit0, int64_t nb, const double* ub, double* part, int* heads) {
    if (!smp.p4) return hipErrorInvalidValue;          // the FP32 element adjoints read the samples' FP32 copy
    return adj_splats_k<f32>(s, pops, a, thetao, soa, npix, N, smp, sp, sadj, ns, knots, ktimes, times, nt, nus, nitems, rec, bar, aflg, unit0, nb,
                             ub, part, heads);
}
hipError_t launch_unit_bounds(hipStream_t s, int64_t npix, int N, Samples smp, const double* times, int nt, int64_t nitems,
                              const uint8_t* aflg, double* ub) {
    hipLaunchKernelGGL(ks_unit_bounds, dim3((unsigned)adj_units(nitems, N)), dim3(64), 0, s, npix, N, smp, times, nt, nitems, aflg, ub);
    return hipGetLastError();
}
hipError_t launch_adj_reduce(hipStream_t s, const double* part, int64_t nunits, int ns, int accumulate, double* dparams) {
    hipLaunchKernelGGL(ks_adj_reduce, dim3((unsigned)(ns * ADJ_ROWS)), dim3(64), 0, s, part, nunits, ns, accumulate, dparams);
    return hipGetLastError();
}

}  // namespace ksd
