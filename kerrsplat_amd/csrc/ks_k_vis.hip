// Visibilities of rendered Stokes images (visibilities.jl:21-40, device_visibilities.jl:53-111).
#include "ks_kernels.hpp"

namespace ksd {

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

hipError_t launch_vis_direct(hipStream_t s, const double* img, const int* perm, int64_t npix, int nalpha, int nbeta, int nframes,
                             double psize, double scale, const double* u, const double* v, const double* tap, int nuv, double* V) {
    hipLaunchKernelGGL(ks_vis_direct, dim3(grid_for((int64_t)nuv * nframes, 64)), dim3(64), 0, s, img, perm, npix, nalpha, nbeta, nframes,
                       psize, scale, u, v, tap, nuv, V);
    return hipGetLastError();
}

}  // namespace ksd
