// Geometry kernels: K0 root cases, the stable 4/2/0 case partition, K1 per-pixel constants (sorted
// structure of arrays), the stored-sample march with anchor segments in parallel, the half-orbit
// cutoff, and the sorted -> screen scatter.
// Reference: pixel_constants.jl:106-176, direct_march.jl:27-124, quadrature.jl:231-367, cache.jl:153-296.
#include "ks_kernels.hpp"

namespace ksd {

__global__ void ks_root_case(double a, double thetao, const double* __restrict__ al, const double* __restrict__ be, int64_t npix,
                             int8_t* __restrict__ numreals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    numreals[i] = (int8_t)root_case_count(a, thetao, al[i], be[i]);
}

// per block: counts of the three cases
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

// ---- stored samples (march!, cache.jl:182-190) ------------------------------------------------------
struct StoreF {
    Samples S; int64_t j, npix; double a, rp, rm;
    KS_DEV void operator()(int k, const Samp& s, double) {
        const int64_t o = (int64_t)(k - 1) * npix + j;
        S.t[o] = s.t; S.r[o] = s.r; S.th[o] = s.th; S.ph[o] = s.ph;
        S.fl[o] = (uint8_t)((s.ok ? 1 : 0) | (s.nur ? 2 : 0) | (s.nuth ? 4 : 0));
        double x, y, z;
        quasi_cartesian_ks(a, rp, rm, s.r, s.th, s.ph, x, y, z);
        S.x[o] = x; S.y[o] = y; S.z[o] = z;
    }
};
// Work items (segment m, pixel j), pixel fastest. Recurrence: the pixel's anchor segments (march_range, its bulk
// in the scalar B: double, or f32 between FP64 anchors); Direct: the closed forms at every sample of chunk m (SEG_MIN
// samples). Anchor residuals per segment res[2 (m npix + j) + {0, 1}] (zero for Direct and past the pixel's last
// segment).
constexpr double FP32_MARCH_RMIN = 1.2;      // x r+
constexpr double FP32_MARCH_RMAX = 20.0;     // M
constexpr double FP32_MARCH_KGAP = 1e-3;     // 1 - k, the radial modulus' distance from the critical curve
template <int MARCHER, class B>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(KS_WAVES_MARCH))) ks_march_store(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                      int nseg_max, Samples smp, double* __restrict__ res) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= npix * nseg_max) return;
    const int64_t j = tid % npix;
    const int m = (int)(tid / npix);
    const Spin S = make_spin(a);
    const PixC p = load_pixel(soa, npix, j);
    StoreF f{smp, j, npix, a, S.rp, S.rm};
    double r1 = 0.0, r2 = 0.0;
    const int L = MARCHER == 0 ? SEG_MIN : segment_length(M, 1.0 - p.k);
    const int kbeg = m * L + 1;
    if (kbeg <= N) {
        const int kend = kbeg + L - 1 < N ? kbeg + L - 1 : N;
        if constexpr (MARCHER == 0) {
            const VisC vis = vis_constants(p, a, thetao);
            const double dtau = p.tau_total / (double)(N + 1);
            for (int k = kbeg; k <= kend; ++k) f(k, direct_sample(p, S, vis, k * dtau), dtau);
        } else if constexpr (std::is_same<B, double>::value) {
            march_range<double>(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
        } else {
            // the FP32 bulk where it keeps its digits: the segment's radii (closed forms at its ends, and the radial
            // turning point when it lies inside) within FP32_MARCH_RMIN r+ .. FP32_MARCH_RMAX, off the critical curve;
            // FP64 elsewhere (the Moebius form of r and the desingularised rates cancel at large r, the logarithmic
            // pieces near the horizon)
            const double dtau = p.tau_total / (double)(N + 1);
            const double ta = kbeg * dtau, tb = kend * dtau;
            const JS ja = radial_state(p, ta), jb = radial_state(p, tb);
            double rlo = dmin(radius_jac(p, ja.sn, ja.cn), radius_jac(p, jb.sn, jb.cn));
            const double rhi = dmax(radius_jac(p, ja.sn, ja.cn), radius_jac(p, jb.sn, jb.cn));
            if (p.rcase == 2 && ta < p.fo && p.fo < tb) rlo = dmin(rlo, p.r4);
            const bool fp32 = rlo > FP32_MARCH_RMIN * S.rp && rhi < FP32_MARCH_RMAX && 1.0 - p.k > FP32_MARCH_KGAP && tb <= p.tau_h;
            if (fp32) march_range<B>(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
            else march_range<double>(p, S, thetao, N, M, kbeg, kend, f, r1, r2);
        }
    }
    res[2 * (m * npix + j)] = r1;
    res[2 * (m * npix + j) + 1] = r2;
}
// the FP32 copy of the sample positions and times the FP32 element adjoints read: one 16-byte load per sample
__global__ void ks_samples_f32(int64_t n, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                               const double* __restrict__ t, float4* __restrict__ p4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p4[i] = make_float4((float)x[i], (float)y[i], (float)z[i], (float)t[i]);
}
hipError_t launch_samples_f32(hipStream_t s, int64_t n, const double* x, const double* y, const double* z, const double* t, float4* p4) {
    hipLaunchKernelGGL(ks_samples_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x, y, z, t, p4);
    return hipGetLastError();
}

// per pixel: the largest residual over its segments (frame 0 of a render, or a stored march)
__global__ void ks_res_reduce(int64_t npix, int nseg_max, const double* __restrict__ res, double* __restrict__ rt, double* __restrict__ rp) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    double r1 = 0.0, r2 = 0.0;
    for (int m = 0; m < nseg_max; ++m) { r1 = dmax(r1, res[2 * (m * npix + j)]); r2 = dmax(r2, res[2 * (m * npix + j) + 1]); }
    rt[j] = r1; rp[j] = r2;
}

// from the march (fused caches) or from the stored samples (smp.r != nullptr)
__global__ void __launch_bounds__(128) ks_wind_cutoff(double a, double thetao, const double* __restrict__ soa, int64_t npix, int N, int M,
                                                      int nmax, double slab, Samples smp, int* __restrict__ kstop) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    Spin S = make_spin(a);
    WindF f{S.rp, slab, nmax, N + 1, NAN, 0, false, false};
    if (smp.r) {
        for (int k = 1; k <= N && f.kstop > N; ++k) {
            const int64_t o = (int64_t)(k - 1) * npix + j;
            const uint8_t fl = smp.fl[o];
            Samp s{0.0, smp.r[o], smp.th[o], 0.0, (fl & 2) != 0, (fl & 4) != 0, (fl & 1) != 0};
            f(k, s, 0.0);
        }
    } else {
        PixC p = load_pixel(soa, npix, j);
        double r1, r2;
        march_ray(p, S, thetao, N, M, f, r1, r2);
    }
    kstop[j] = f.kstop;
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

// ---- launchers ----
hipError_t launch_root_case(hipStream_t s, double a, double thetao, const double* al, const double* be, int64_t npix, int8_t* numreals) {
    hipLaunchKernelGGL(ks_root_case, dim3(grid_for(npix, 256)), dim3(256), 0, s, a, thetao, al, be, npix, numreals);
    return hipGetLastError();
}
hipError_t launch_case_count(hipStream_t s, int nblocks, const int8_t* numreals, int64_t npix, int* bcount) {
    hipLaunchKernelGGL(ks_case_count, dim3(nblocks), dim3(PART_BLOCK), 0, s, numreals, npix, bcount);
    return hipGetLastError();
}
hipError_t launch_case_scatter(hipStream_t s, int nblocks, const int8_t* numreals, int64_t npix, const int* boff, int* perm) {
    hipLaunchKernelGGL(ks_case_scatter, dim3(nblocks), dim3(PART_BLOCK), 0, s, numreals, npix, boff, perm);
    return hipGetLastError();
}
hipError_t launch_pixel_constants(hipStream_t s, double a, double thetao, const double* al, const double* be, const int* perm, int64_t npix,
                                  double* soa) {
    hipLaunchKernelGGL(ks_pixel_constants, dim3(grid_for(npix, 128)), dim3(128), 0, s, a, thetao, al, be, perm, npix, soa);
    return hipGetLastError();
}
hipError_t launch_march_store(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, int marcher, int nseg_max,
                              Samples smp, double* res, int fp32) {
    const dim3 grid(grid_for(npix * nseg_max, 128)), block(128);
    if (marcher == 0) hipLaunchKernelGGL((ks_march_store<0, double>), grid, block, 0, s, a, thetao, soa, npix, N, M, nseg_max, smp, res);
    else if (fp32) hipLaunchKernelGGL((ks_march_store<1, f32>), grid, block, 0, s, a, thetao, soa, npix, N, M, nseg_max, smp, res);
    else hipLaunchKernelGGL((ks_march_store<1, double>), grid, block, 0, s, a, thetao, soa, npix, N, M, nseg_max, smp, res);
    return hipGetLastError();
}
hipError_t launch_res_reduce(hipStream_t s, int64_t npix, int nseg_max, const double* res, double* rt, double* rp) {
    hipLaunchKernelGGL(ks_res_reduce, dim3(grid_for(npix, 256)), dim3(256), 0, s, npix, nseg_max, res, rt, rp);
    return hipGetLastError();
}
hipError_t launch_wind_cutoff(hipStream_t s, double a, double thetao, const double* soa, int64_t npix, int N, int M, int nmax, double slab,
                              Samples smp, int* kstop) {
    hipLaunchKernelGGL(ks_wind_cutoff, dim3(grid_for(npix, 128)), dim3(128), 0, s, a, thetao, soa, npix, N, M, nmax, slab, smp, kstop);
    return hipGetLastError();
}
hipError_t launch_unsort4(hipStream_t s, const double* src, const int* perm, int64_t npix, int64_t nframes, double* dst) {
    hipLaunchKernelGGL(ks_unsort4, dim3(grid_for(npix * nframes, 256)), dim3(256), 0, s, src, perm, npix, nframes, dst);
    return hipGetLastError();
}
hipError_t launch_unsort1(hipStream_t s, const double* src, const int* perm, int64_t npix, double* dst) {
    hipLaunchKernelGGL(ks_unsort1, dim3(grid_for(npix, 256)), dim3(256), 0, s, src, perm, npix, dst);
    return hipGetLastError();
}

}  // namespace ksd
