// Direct Fourier transform of Stokes images onto (u, v) points and its adjoint seed, for the CPU
// oracle (test infrastructure; see ks_scalar.hpp). Restates src/Fit/visibilities.jl:21-40 and
// src/Fit/device_visibilities.jl:53-88 (EHT sign convention):
//   V_k = tap_k sum_pixels I_pixel (psize^2 / Jy) exp(+2 pi i (u_k l + v_k m)),
//   l = -(i - (n_alpha + 1)/2) psize (east = -alpha), m = (j - (n_beta + 1)/2) psize, i, j 1-based;
//   seed_pixel = sum_k Re[w_k conj(dV_k/dI_pixel)] = sum_k Re[w_k tap_k scale exp(-i phase)].
// Images are n_alpha x n_beta x 4 per frame in screen (column-major) order.
#pragma once
#include "ks_transfer.hpp"

namespace kso {

inline void visibilities(int na, int nb, int nf, const double* images, double psize, int nuv, const double* u, const double* v,
                         const double* tap, double* Vre, double* Vim) {
    double scale = psize * psize / JY;
    int npix = na * nb;
#pragma omp parallel for schedule(static) collapse(2)
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < nuv; ++k) {
            double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
            for (int j = 0; j < nb; ++j) {
                double m = ((j + 1) - (nb + 1) / 2.0) * psize;
                for (int i = 0; i < na; ++i) {
                    double l = -((i + 1) - (na + 1) / 2.0) * psize;
                    double ph = 2.0 * KS_PI * (u[k] * l + v[k] * m);
                    double s = std::sin(ph), c = std::cos(ph);
                    const double* I = images + 4 * ((size_t)f * npix + i + (size_t)na * j);
                    for (int q = 0; q < 4; ++q) { re[q] += I[q] * scale * c; im[q] += I[q] * scale * s; }
                }
            }
            double t = tap ? tap[k] : 1.0;
            for (int q = 0; q < 4; ++q) {
                Vre[4 * ((size_t)f * nuv + k) + q] = re[q] * t;
                Vim[4 * ((size_t)f * nuv + k) + q] = im[q] * t;
            }
        }
}

inline void visibility_seed(int na, int nb, int nf, double psize, int nuv, const double* u, const double* v, const double* tap,
                            const double* Wre, const double* Wim, double* seed) {
    double scale = psize * psize / JY;
    int npix = na * nb;
#pragma omp parallel for schedule(static) collapse(2)
    for (int f = 0; f < nf; ++f)
        for (int p = 0; p < npix; ++p) {
            int i = p % na, j = p / na;
            double l = -((i + 1) - (na + 1) / 2.0) * psize, m = ((j + 1) - (nb + 1) / 2.0) * psize;
            double acc[4] = {0, 0, 0, 0};
            for (int k = 0; k < nuv; ++k) {
                double ph = 2.0 * KS_PI * (u[k] * l + v[k] * m);
                double s = std::sin(ph), c = std::cos(ph);
                double t = (tap ? tap[k] : 1.0) * scale;
                for (int q = 0; q < 4; ++q) {
                    double wr = Wre[4 * ((size_t)f * nuv + k) + q] * t, wi = Wim[4 * ((size_t)f * nuv + k) + q] * t;
                    acc[q] += wr * c + wi * s;      // Re[(wr + i wi)(c - i s)]
                }
            }
            for (int q = 0; q < 4; ++q) seed[4 * ((size_t)f * npix + p) + q] = acc[q];
        }
}

}  // namespace kso
