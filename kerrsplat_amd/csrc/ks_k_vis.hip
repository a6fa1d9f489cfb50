// Visibilities of rendered Stokes images, their χ² against data and the adjoint seeds
// (visibilities.jl:21-40, device_visibilities.jl:53-111; the image χ² of _frame_seed!, Fit.jl:537-577).
//
// On Krang's pixel grid the phase of pixel (i, j) at baseline k separates,
//   2 pi (u_k l_i + v_k m_j) = 2 pi u_k l_i + 2 pi v_k m_j,
// so V_k = tap_k scale sum_j By[k, j] sum_i Ax[k, i] I[i, j] with Ax = e^{2 pi i u l}, By = e^{2 pi i v m}.
// The inner sum over i is a real matrix product with the phase table stacked as [cos; sin]
// (2 nuv x n_alpha) against the image as an n_alpha x 4 n_beta matrix; the adjoint seed
// Re sum_k w_k conj(Ax By) is the transposed product of the same table with the weights rotated by
// conj(By). Both products run through one FP64 GEMM on the matrix cores (v_mfma_f64_16x16x4f64);
// everything else is O(nuv (n_alpha + n_beta)).
#include "ks_kernels.hpp"

namespace ksd {

// ---- phase tables: A2[k][i] = cos(2 pi u_k l_i), A2[nuv + k][i] = sin(...); B2 likewise for v_k m_j ---------------
// l_i = -((i + 1) - (n_alpha + 1) / 2) psize (east = -alpha), m_j = ((j + 1) - (n_beta + 1) / 2) psize
__global__ void ks_vis_tables(int nuv, int na, int nb, double psize, const double* __restrict__ u, const double* __restrict__ v,
                              double* __restrict__ A2, double* __restrict__ B2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int row = na + nb;
    if (t >= (int64_t)nuv * row) return;
    const int k = (int)(t / row), q = (int)(t % row);
    double s, c;
    if (q < na) {
        const double l = -((q + 1) - (na + 1) * 0.5) * psize;
        sincos(2.0 * PI * (u[k] * l), &s, &c);
        A2[(int64_t)k * na + q] = c;
        A2[(int64_t)(nuv + k) * na + q] = s;
    } else {
        const int j = q - na;
        const double m = ((j + 1) - (nb + 1) * 0.5) * psize;
        sincos(2.0 * PI * (v[k] * m), &s, &c);
        B2[(int64_t)k * nb + j] = c;
        B2[(int64_t)(nuv + k) * nb + j] = s;
    }
}

// ---- the image as n_alpha x 4 n_beta row-major matrices, one per frame, from the sorted cube ----------------------
// Im[f][i][4 j + q] = out[4 (s + npix f) + q] with perm[s] = i + n_alpha j
__global__ void ks_vis_gather(const double* __restrict__ out, const int* __restrict__ perm, int64_t npix, int na, int nb, int64_t f0, int nf,
                              double* __restrict__ Im) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npix * nf) return;
    const int64_t s = t % npix, f = t / npix;
    const int p = perm[s];
    const int i = p % na, j = p / na;
    const double* src = out + 4 * (s + npix * (f0 + f));
    double* dst = Im + (int64_t)npix * 4 * f + (int64_t)i * 4 * nb + 4 * j;
    for (int q = 0; q < 4; ++q) dst[q] = src[q];
}

// ---- FP64 GEMM on the matrix cores: C[M x N] = op(A) B, row-major; A is M x K, or K x M with TA; batched over z ----
// 64 x 64 tile per 256-thread block, K in steps of 16 through LDS (the next step's operands loaded into registers
// during this step's products). Wavefront w owns rows 16 w .. 16 w + 15
// and the four 16 x 16 output tiles along N, each a v_mfma_f64_16x16x4f64 accumulator; per step of 4 in K
// lane l supplies A[16 w + l % 16][k + l / 16] and B[k + l / 16][16 t + l % 16], and holds
// D[4 r + l / 16][l % 16] in its register r (the instruction's lane layout, checked on the device).
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64, GK = 16;
template <bool TA>
__global__ void __launch_bounds__(256) ks_dgemm(int M, int N, int K, const double* __restrict__ A, int64_t lda, int64_t sA,
                                                const double* __restrict__ B, int64_t ldb, int64_t sB, double* __restrict__ C, int64_t ldc,
                                                int64_t sC) {
    __shared__ double As[GK][GT + 1];
    __shared__ double Bs[GK][GT + 1];
    A += sA * blockIdx.z; B += sB * blockIdx.z; C += sC * blockIdx.z;
    const int tid = threadIdx.x, w = tid / 64, l = tid % 64;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    dbl4 acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    // each thread moves EPT elements of A and EPT of B per step of K (GT x GK = EPT x 256); the next step's elements
    // are loaded into registers while the matrix cores work on this step's tiles
    constexpr int EPT = GT * GK / 256;
    static_assert(EPT * 256 == GT * GK, "whole elements of each operand per thread and step");
    double ra[EPT], rb[EPT];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = tid + 256 * u;
            const int mm = TA ? e % GT : e / GK, kk = TA ? e / GT : e % GK;
            const int gm = m0 + mm, gk = k0 + kk;
            ra[u] = (gm < M && gk < K) ? (TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk]) : 0.0;
            const int gk2 = k0 + e / GT, gn = n0 + e % GT;
            rb[u] = (gk2 < K && gn < N) ? B[(int64_t)gk2 * ldb + gn] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = tid + 256 * u;
            const int mm = TA ? e % GT : e / GK, kk = TA ? e / GT : e % GK;
            As[kk][mm] = ra[u];
            Bs[e / GT][e % GT] = rb[u];
        }
        __syncthreads();
        if (k0 + GK < K) load(k0 + GK);
#pragma unroll
        for (int k4 = 0; k4 < GK; k4 += 4) {
            const double a = As[k4 + l / 16][16 * w + l % 16];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[k4 + l / 16][16 * t + l % 16], acc[t], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int t = 0; t < 4; ++t) {
        const int gn = n0 + 16 * t + l % 16;
        if (gn >= N) continue;
        for (int r = 0; r < 4; ++r) {
            const int gm = m0 + 16 * w + 4 * r + l / 16;
            if (gm < M) C[(int64_t)gm * ldc + gn] = acc[t][r];
        }
    }
}

// C[i] = sum over the nsplit partial products P[z][i] in order (the split-K GEMM's fixed-order reduction)
__global__ void ks_sum_partials(const double* __restrict__ P, int nsplit, int64_t n, double* __restrict__ C) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += P[(int64_t)z * n + i];
    C[i] = s;
}

// ---- V[f][k][q] = tap_k scale sum_j (Cr + i Ci)[k][4 j + q] (Br + i Bi)[k][j]; V as (re, im) pairs ------------------
// C2 per frame: 2 nuv x 4 n_beta (rows k: real part, rows nuv + k: imaginary part)
__global__ void ks_vis_combine(int nuv, int nb, int nf, const double* __restrict__ C2, const double* __restrict__ B2,
                               const double* __restrict__ tap, double scale, double* __restrict__ V) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nuv * 4 * nf) return;
    const int q = (int)(t % 4), k = (int)((t / 4) % nuv);
    const int64_t f = t / (4 * (int64_t)nuv);
    const double* Cr = C2 + (int64_t)f * 2 * nuv * 4 * nb + (int64_t)k * 4 * nb + q;
    const double* Ci = Cr + (int64_t)nuv * 4 * nb;
    const double* Br = B2 + (int64_t)k * nb;
    const double* Bi = Br + (int64_t)nuv * nb;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < nb; ++j) {
        const double cr = Cr[4 * j], ci = Ci[4 * j], br = Br[j], bi = Bi[j];
        re += cr * br - ci * bi;
        im += cr * bi + ci * br;
    }
    const double w = (tap ? tap[k] : 1.0) * scale;
    V[2 * t] = re * w;
    V[2 * t + 1] = im * w;
}

// ---- residuals r = (g V - d) / sigma per Stokes parameter, chi2 per baseline, and the weights rotated for the seed --
// w_k = 2 conj(g_k) r_k / sigma_k (device_visibilities.jl:106), times tap_k scale (dV/dI); the seed product
// then takes E2[k][4 j + q] = Re(w_kq conj(By[k][j])), E2[nuv + k][4 j + q] = Im(...).
// V, vis: nuv x 4 complex (re, im); sigma: nuv x 4; gains: nuv complex or null.
__global__ void ks_vis_residual(int nuv, const double* __restrict__ V, const double* __restrict__ vis, const double* __restrict__ sigma,
                               const double* __restrict__ gains, const double* __restrict__ tap, double scale, double* __restrict__ wq,
                               double* __restrict__ chi_k) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nuv) return;
    const double gr = gains ? gains[2 * k] : 1.0, gi = gains ? gains[2 * k + 1] : 0.0;
    const double t = (tap ? tap[k] : 1.0) * scale;
    double chi = 0.0;
    for (int q = 0; q < 4; ++q) {
        const int64_t o = 4 * (int64_t)k + q;
        const double vr = V[2 * o], vi = V[2 * o + 1];
        const double is = 1.0 / sigma[o];
        const double rr = (gr * vr - gi * vi - vis[2 * o]) * is, ri = (gr * vi + gi * vr - vis[2 * o + 1]) * is;
        chi += rr * rr + ri * ri;
        // 2 conj(g) r / sigma
        wq[2 * o] = 2.0 * (gr * rr + gi * ri) * is * t;
        wq[2 * o + 1] = 2.0 * (gr * ri - gi * rr) * is * t;
    }
    chi_k[k] = chi;
}
__global__ void ks_vis_rotate(int nuv, int nb, const double* __restrict__ wq, const double* __restrict__ B2, double* __restrict__ E2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nuv * nb) return;
    const int k = (int)(t / nb), j = (int)(t % nb);
    const double br = B2[(int64_t)k * nb + j], bi = B2[(int64_t)(nuv + k) * nb + j];
    double* er = E2 + (int64_t)k * 4 * nb + 4 * j;
    double* ei = er + (int64_t)nuv * 4 * nb;
    for (int q = 0; q < 4; ++q) {
        const double wr = wq[2 * (4 * (int64_t)k + q)], wi = wq[2 * (4 * (int64_t)k + q) + 1];
        er[q] = wr * br + wi * bi;      // w conj(b)
        ei[q] = wi * br - wr * bi;
    }
}

// seed (n_alpha x 4 n_beta, screen) -> the dual sweep's seed slot of frame f in sorted order (added to it with add)
__global__ void ks_vis_scatter_seed(const double* __restrict__ Sd, const int* __restrict__ perm, int64_t npix, int na, int nb, int64_t f,
                                    int add, double* __restrict__ ads) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npix) return;
    const int p = perm[s];
    const int i = p % na, j = p / na;
    const double* src = Sd + (int64_t)i * 4 * nb + 4 * j;
    double* dst = ads + 4 * (s + npix * f);
    for (int q = 0; q < 4; ++q) dst[q] = add ? dst[q] + src[q] : src[q];
}
// weights w (nuv x 4 complex, the derivative of a loss with respect to the model visibilities) times
// tap_k scale: the seed_kernel! input of an arbitrary visibility-domain loss
__global__ void ks_vis_scale_weights(int nuv, const double* __restrict__ w, const double* __restrict__ tap, double scale, double* __restrict__ wq) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nuv) return;
    const double t = (tap ? tap[k] : 1.0) * scale;
    for (int q = 0; q < 8; ++q) wq[8 * (int64_t)k + q] = w[8 * (int64_t)k + q] * t;
}

// ---- closure phases and log closure amplitudes of one frame (chi2_closures, visibilities.jl:141-199) ------------------
// The Stokes I visibility of 1-based signed baseline index k (negative: the conjugate), as (re, im).
KS_DEV void signed_vis(const double* __restrict__ V, int k, double& re, double& im) {
    const int64_t o = 8 * (int64_t)(k < 0 ? -k - 1 : k - 1);
    re = V[o]; im = k < 0 ? -V[o + 1] : V[o + 1];
}
// One thread per closure term: triangles first (the bispectrum's phase against the data, the difference taken to
// the nearest multiple of 2 pi, rem(., 2 pi, RoundNearest)), then quadrangles (ln |V1||V2| / (|V3||V4|) against the
// data). chi_t: the term's chi2; coef: d chi2 / d (phase or log amplitude).
__global__ void ks_closure_terms(int ntri, const int* __restrict__ tri, const double* __restrict__ cp_data, const double* __restrict__ cp_sig,
                                 int nquad, const int* __restrict__ quad, const double* __restrict__ la_data, const double* __restrict__ la_sig,
                                 const double* __restrict__ V, double* __restrict__ chi_t, double* __restrict__ coef) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri + nquad) return;
    if (t < ntri) {
        double br = 1.0, bi = 0.0;
        for (int l = 0; l < 3; ++l) {
            double zr, zi;
            signed_vis(V, tri[3 * t + l], zr, zi);
            const double nr = br * zr - bi * zi, ni = br * zi + bi * zr;
            br = nr; bi = ni;
        }
        const double d = remainder(atan2(bi, br) - cp_data[t], 2.0 * PI);
        const double is = 1.0 / cp_sig[t];
        chi_t[t] = d * d * is * is;
        coef[t] = 2.0 * d * is * is;
    } else {
        const int q = t - ntri;
        double la = 0.0;
        for (int l = 0; l < 4; ++l) {
            double zr, zi;
            signed_vis(V, quad[4 * q + l], zr, zi);
            const double lz = 0.5 * log(zr * zr + zi * zi);
            la += l < 2 ? lz : -lz;
        }
        const double r = (la - la_data[q]) / la_sig[q];
        chi_t[t] = r * r;
        coef[t] = 2.0 * r / la_sig[q];
    }
}
// One thread per baseline k: w = d chi2 / d Re V + i d chi2 / d Im V of its Stokes I visibility (Q, U, V zero), summed
// over the terms that use it in a fixed order (the lists legs[off[k] .. off[k + 1]): entry = term index, with the leg's
// role in bit 30 (set: a conjugated triangle leg or a denominator amplitude)). d arg V = Im(dV / V), d ln |V| = Re(dV / V):
// w += c i V / |V|^2 for a triangle leg (negated when conjugated), c V / |V|^2 for a numerator amplitude (negated in
// the denominator).
__global__ void ks_closure_weights(int nuv, int ntri, const int* __restrict__ off, const int* __restrict__ legs, const double* __restrict__ coef,
                                   const double* __restrict__ V, double* __restrict__ w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nuv) return;
    const double zr = V[8 * (int64_t)k], zi = V[8 * (int64_t)k + 1];
    const double n2 = zr * zr + zi * zi;
    double wr = 0.0, wi = 0.0;
    if (n2 > 0.0)
        for (int e = off[k]; e < off[k + 1]; ++e) {
            const int t = legs[e] & 0x3fffffff;
            const double s = (legs[e] & 0x40000000) ? -1.0 : 1.0;
            const double c = s * coef[t] / n2;
            if (t < ntri) { wr += -c * zi; wi += c * zr; }     // c i z
            else { wr += c * zr; wi += c * zi; }               // c z
        }
    double* o = w + 8 * (int64_t)k;
    o[0] = wr; o[1] = wi;
    for (int q = 2; q < 8; ++q) o[q] = 0.0;
}

// ---- the image χ² of one frame (_frame_seed!, unbinned): r = (I - d) / sigma over masked pixels, seed 2 r / sigma ---
// data, sigma (4 or 4 per pixel) and mask (per pixel, or null) in screen order; image and seed in sorted order.
__global__ void ks_image_residual(const double* __restrict__ out, const int* __restrict__ perm, int64_t npix, int64_t f,
                                  const double* __restrict__ data, const double* __restrict__ sigma, int sigma_per_pixel,
                                  const uint8_t* __restrict__ mask, double* __restrict__ ads, double* __restrict__ chi_p) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npix) return;
    const int p = perm[s];
    const double* img = out + 4 * (s + npix * f);
    double* seed = ads + 4 * (s + npix * f);
    double chi = 0.0;
    if (!mask || mask[p]) {
        const double* sg = sigma + (sigma_per_pixel ? 4 * (int64_t)p : 0);
        for (int q = 0; q < 4; ++q) {
            const double is = 1.0 / sg[q];
            const double r = (img[q] - data[4 * (int64_t)p + q]) * is;
            chi += r * r;
            seed[q] = 2.0 * r * is;
        }
    } else {
        for (int q = 0; q < 4; ++q) seed[q] = 0.0;
    }
    chi_p[s] = chi;
}

// _frame_seed! through a Binning (Fit.jl:555-574), per binned pixel q: the mean of its points' images (the
// points in ascending screen order, as bin() adds them, camera.jl:91-98), the weighted residual, its χ² term
// and the seed 2 r / sigma / count that each of its points receives. iperm: screen index -> sorted slot.
__global__ void ks_bin_residual(const double* __restrict__ out, const int* __restrict__ iperm, int64_t npix, int64_t f,
                                const int* __restrict__ off, const int* __restrict__ pts, int64_t nbins, const double* __restrict__ data,
                                const double* __restrict__ sigma, int sigma_per_pixel, const uint8_t* __restrict__ mask,
                                double* __restrict__ pseed, double* __restrict__ chi_p) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nbins) return;
    const int b = off[q], e = off[q + 1];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int m = b; m < e; ++m) {
        const double* img = out + 4 * ((int64_t)iperm[pts[m]] + npix * f);
        for (int k = 0; k < 4; ++k) acc[k] += img[k];
    }
    const double cnt = (double)(e - b);
    double chi = 0.0;
    if (!mask || mask[q]) {
        const double* sg = sigma + (sigma_per_pixel ? 4 * q : 0);
        for (int k = 0; k < 4; ++k) {
            const double r = (acc[k] / cnt - data[4 * q + k]) / sg[k];
            chi += r * r;
            pseed[4 * q + k] = 2.0 * r / sg[k] / cnt;
        }
    } else {
        for (int k = 0; k < 4; ++k) pseed[4 * q + k] = 0.0;
    }
    chi_p[q] = chi;
}

// each sorted slot takes the seed of the binned pixel its point belongs to
__global__ void ks_bin_seed(const int* __restrict__ perm, const int* __restrict__ pix, int64_t npix, int64_t f,
                            const double* __restrict__ pseed, double* __restrict__ ads) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npix) return;
    const int64_t q = pix[perm[s]];
    for (int k = 0; k < 4; ++k) ads[4 * (s + npix * f) + k] = pseed[4 * q + k];
}

// ---- deterministic sum of n values by one block: strided per-thread sums, then a fixed tree ----------------------
__global__ void __launch_bounds__(256) ks_sum_fixed(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ---- launchers ----------------------------------------------------------------------------------------------------
hipError_t launch_vis_tables(hipStream_t s, int nuv, int na, int nb, double psize, const double* u, const double* v, double* A2, double* B2) {
    hipLaunchKernelGGL(ks_vis_tables, dim3(grid_for((int64_t)nuv * (na + nb), 256)), dim3(256), 0, s, nuv, na, nb, psize, u, v, A2, B2);
    return hipGetLastError();
}
hipError_t launch_vis_gather(hipStream_t s, const double* out, const int* perm, int64_t npix, int na, int nb, int64_t f0, int nf, double* Im) {
    hipLaunchKernelGGL(ks_vis_gather, dim3(grid_for(npix * nf, 256)), dim3(256), 0, s, out, perm, npix, na, nb, f0, nf, Im);
    return hipGetLastError();
}
hipError_t launch_dgemm(hipStream_t s, bool transA, int M, int N, int K, const double* A, int64_t lda, int64_t sA, const double* B,
                        int64_t ldb, int64_t sB, double* C, int64_t ldc, int64_t sC, int batch) {
    const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, batch);
    if (transA) hipLaunchKernelGGL(ks_dgemm<true>, grid, dim3(256), 0, s, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
    else hipLaunchKernelGGL(ks_dgemm<false>, grid, dim3(256), 0, s, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
    return hipGetLastError();
}
// C = A^T B with A stored K x M, K cut into nsplit equal ranges (nsplit halved until it divides K) whose
// products land in P (nsplit x M x N) in one launch and are summed in order into C: the split gives a
// short-M, long-K product enough blocks to fill the chip
hipError_t launch_dgemm_tn_splitk(hipStream_t s, int M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb,
                                  double* C, int nsplit, double* P) {
    while (nsplit > 1 && K % nsplit) nsplit /= 2;
    const int kc = K / nsplit;
    const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, nsplit);
    hipLaunchKernelGGL(ks_dgemm<true>, grid, dim3(256), 0, s, M, N, kc, A, lda, (int64_t)kc * lda, B, ldb, (int64_t)kc * ldb, P, (int64_t)N,
                       (int64_t)M * N);
    hipLaunchKernelGGL(ks_sum_partials, dim3(grid_for((int64_t)M * N, 256)), dim3(256), 0, s, P, nsplit, (int64_t)M * N, C);
    return hipGetLastError();
}
hipError_t launch_vis_combine(hipStream_t s, int nuv, int nb, int nf, const double* C2, const double* B2, const double* tap, double scale, double* V) {
    hipLaunchKernelGGL(ks_vis_combine, dim3(grid_for((int64_t)nuv * 4 * nf, 256)), dim3(256), 0, s, nuv, nb, nf, C2, B2, tap, scale, V);
    return hipGetLastError();
}
hipError_t launch_vis_residual(hipStream_t s, int nuv, const double* V, const double* vis, const double* sigma, const double* gains,
                               const double* tap, double scale, double* wq, double* chi_k) {
    hipLaunchKernelGGL(ks_vis_residual, dim3(grid_for(nuv, 256)), dim3(256), 0, s, nuv, V, vis, sigma, gains, tap, scale, wq, chi_k);
    return hipGetLastError();
}
hipError_t launch_vis_rotate(hipStream_t s, int nuv, int nb, const double* wq, const double* B2, double* E2) {
    hipLaunchKernelGGL(ks_vis_rotate, dim3(grid_for((int64_t)nuv * nb, 256)), dim3(256), 0, s, nuv, nb, wq, B2, E2);
    return hipGetLastError();
}
hipError_t launch_vis_scatter_seed(hipStream_t s, const double* Sd, const int* perm, int64_t npix, int na, int nb, int64_t f, int add, double* ads) {
    hipLaunchKernelGGL(ks_vis_scatter_seed, dim3(grid_for(npix, 256)), dim3(256), 0, s, Sd, perm, npix, na, nb, f, add, ads);
    return hipGetLastError();
}
hipError_t launch_vis_scale_weights(hipStream_t s, int nuv, const double* w, const double* tap, double scale, double* wq) {
    hipLaunchKernelGGL(ks_vis_scale_weights, dim3(grid_for(nuv, 256)), dim3(256), 0, s, nuv, w, tap, scale, wq);
    return hipGetLastError();
}
hipError_t launch_image_residual(hipStream_t s, const double* out, const int* perm, int64_t npix, int64_t f, const double* data,
                                 const double* sigma, int sigma_per_pixel, const uint8_t* mask, double* ads, double* chi_p) {
    hipLaunchKernelGGL(ks_image_residual, dim3(grid_for(npix, 256)), dim3(256), 0, s, out, perm, npix, f, data, sigma, sigma_per_pixel, mask, ads, chi_p);
    return hipGetLastError();
}
hipError_t launch_closure_terms(hipStream_t s, int ntri, const int* tri, const double* cp_data, const double* cp_sig, int nquad, const int* quad,
                                const double* la_data, const double* la_sig, const double* V, double* chi_t, double* coef) {
    hipLaunchKernelGGL(ks_closure_terms, dim3(grid_for(ntri + nquad, 256)), dim3(256), 0, s, ntri, tri, cp_data, cp_sig, nquad, quad, la_data,
                       la_sig, V, chi_t, coef);
    return hipGetLastError();
}
hipError_t launch_closure_weights(hipStream_t s, int nuv, int ntri, const int* off, const int* legs, const double* coef, const double* V, double* w) {
    hipLaunchKernelGGL(ks_closure_weights, dim3(grid_for(nuv, 256)), dim3(256), 0, s, nuv, ntri, off, legs, coef, V, w);
    return hipGetLastError();
}
hipError_t launch_sum_fixed(hipStream_t s, const double* x, int64_t n, double* out) {
    hipLaunchKernelGGL(ks_sum_fixed, dim3(1), dim3(256), 0, s, x, n, out);
    return hipGetLastError();
}
hipError_t launch_bin_residual(hipStream_t s, const double* out, const int* iperm, int64_t npix, int64_t f, const int* off, const int* pts,
                               int64_t nbins, const double* data, const double* sigma, int sigma_per_pixel, const uint8_t* mask,
                               double* pseed, double* chi_p) {
    hipLaunchKernelGGL(ks_bin_residual, dim3(grid_for(nbins, 256)), dim3(256), 0, s, out, iperm, npix, f, off, pts, nbins, data, sigma,
                       sigma_per_pixel, mask, pseed, chi_p);
    return hipGetLastError();
}
hipError_t launch_bin_seed(hipStream_t s, const int* perm, const int* pix, int64_t npix, int64_t f, const double* pseed, double* ads) {
    hipLaunchKernelGGL(ks_bin_seed, dim3(grid_for(npix, 256)), dim3(256), 0, s, perm, pix, npix, f, pseed, ads);
    return hipGetLastError();
}

}  // namespace ksd
