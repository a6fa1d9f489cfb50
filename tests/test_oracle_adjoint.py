"""The oracle's dual-sweep gradient against central finite differences of its own renderer.

The reference checks its image gradients the same way, against central finite differences of
the image loss (test/test_polarized_splats.jl:95-127). Here the loss is
l = sum_pixels dstokes . I_obs with seeded random dstokes, the step is 1e-6 relative (absolute 1e-6
for parameters near zero), and every one of the 21 rows of every splat is differenced.

Three regimes: the default optically thin parcels, parcels 30 times denser (so absorption and
Faraday rotation change the answer at first order), and the half-orbit truncation (nmax = 0 with a
slab), which drops every sample past the first passage through the equatorial slab.
"""
import math

import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn
from kerrsplat_amd.api import Camera

A, THO = 0.5, math.radians(60.0)
NU = 230e9
N = 128


def _setup(dens_boost, seed):
    cam = Camera((-15, 15), (-15, 15), 10)
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(dens_boost)
    ds = np.random.default_rng(seed).normal(size=(cam.npix, 4))
    return cam, P, ds


@pytest.mark.parametrize("dens_boost,nmax,slab", [(1.0, -1, 0.0), (30.0, -1, 0.0), (30.0, 0, 0.5)])
def test_gradient_vs_central_differences(oracle, dens_boost, nmax, slab):
    cam, P, ds = _setup(dens_boost, 5)
    L = syn.length_unit()
    g, img = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, ds, nmax=nmax, slab=slab)
    ref, _, _ = oracle.polarized_image(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, nmax=nmax, slab=slab)
    assert np.abs(img - ref).max() <= 1e-14 * np.abs(ref).max()

    def loss(Q):
        im, _, _ = oracle.polarized_image(A, THO, cam.alphas, cam.betas, N, Q, 0.0, NU, L, nmax=nmax, slab=slab)
        return float(np.sum(ds * im))

    gmax = np.abs(g).max()
    assert gmax > 0
    worst = 0.0
    for i in range(P.shape[1]):
        for r in range(P.shape[0]):
            h = 1e-6 * max(abs(P[r, i]), 1.0)
            Qp, Qm = P.copy(), P.copy()
            Qp[r, i] += h
            Qm[r, i] -= h
            fd = (loss(Qp) - loss(Qm)) / (2 * h)
            worst = max(worst, abs(fd - g[r, i]) / gmax)
    # central differences at h = 1e-6 carry ~1e-9 relative truncation/rounding noise
    assert worst < 1e-6, worst


def test_gradient_is_linear_in_seed(oracle):
    cam, P, ds = _setup(30.0, 7)
    ds2 = np.random.default_rng(8).normal(size=ds.shape)
    L = syn.length_unit()
    g1, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, ds)
    g2, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, ds2)
    g12, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, 2.0 * ds - 3.0 * ds2)
    assert np.abs(g12 - (2.0 * g1 - 3.0 * g2)).max() <= 1e-12 * np.abs(g12).max()


@pytest.mark.parametrize("nmax,slab", [(-1, 0.0), (0, 0.5)])
def test_analytic_weight_derivative_matches_duals(oracle, nmax, slab):
    """d ln G of the splat weight written out by hand (centre under the pattern rotation, widths,
    normalised quaternion, t0, temporal width, pattern rate) equals the 13-partial dual derivative."""
    cam, P, ds = _setup(30.0, 9)
    P = P.copy()
    P[20, :] = [0.05, -0.03, 0.08]          # a visible pattern rotation
    P[11, :] = np.log([40.0, 80.0, 25.0])   # finite temporal envelopes
    P[10, :] = [5.0, -10.0, 2.0]
    L = syn.length_unit()
    g_dual, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 3.0, NU, L, ds, nmax=nmax, slab=slab)
    g_an, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 3.0, NU, L, ds, nmax=nmax, slab=slab, analytic_weight=True)
    assert np.abs(g_dual).max() > 0
    assert np.abs(g_an - g_dual).max() <= 1e-12 * np.abs(g_dual).max()
