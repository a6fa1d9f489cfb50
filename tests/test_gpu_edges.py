"""Edge cases of the device pipeline against the oracle: an empty splat set, a sample count that is a
multiple of neither the 32-sample chunk nor the anchor segments (odd N puts the radial turning point of
scattering rays on a sample: tau_total = 2 fo, dtau = tau_total / (N + 1)), screens that fill neither a
wavefront nor the 8 x 8 tiles of the wavefront order, and a one-pixel camera. Screens with an odd number
of columns hold rays with alpha = 0 exactly (zero angular momentum, through the spin axis), where the
reference's polar term is 0 x infinity once u+ rounds to 1: those rays' phi is NaN in both implementations
or in one of them, as rounding decides, and they are compared for finiteness only. The bars are those of the other
parity tests: images within 1e-10 of the peak Stokes I, gradients within 1e-9 of the largest entry
(FP64); the mixed path within 1e-4 and 1e-3 of the FP64 path.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

A, TH = 0.9, math.radians(45.0)


def cache_of(cam, N, path):
    if path == "stored":
        c = K.GeodesicCache(cam, N, store_samples=True)
        K.regenerate(c, A, TH, marcher=K.Recurrence(64))
    else:
        c = K.GeodesicCache(cam, N, store_samples=False)
        K.regenerate(c, A, TH, marcher=K.Fused(64))
    return c


def test_empty_splat_set():
    cam = K.Camera((-10.0, 10.0), (-10.0, 10.0), (6, 5))
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)[:, :0]
    for path in ("stored", "fused"):
        c = cache_of(cam, 64, path)
        assert np.all(K.polarized_image(c, P, 0.0, 230e9, L) == 0.0)
    c = cache_of(cam, 64, "stored")
    g, cube = K.polarized_gradient(c, P, [0.0], [230e9], L, np.ones((6, 5, 1, 1, 4)))
    assert g.shape == (21, 0) and np.all(cube == 0.0)
    lists = K.ray_lists(c, K.Model(P), 0.0)
    assert lists.capacity == 1 and np.all(lists.count == 0) and np.all(lists.ids == -1)


# 10 x 11 at N = 161: its most different pixel (alpha = -7, beta = 1.8) agrees to 2.9e-10 of the peak while its
# samples agree to 1e-11 (t) and 1e-13 (r, theta), so the difference arises in the transfer, where device and
# oracle round differently (the step as a polynomial in K' against the oracle's form); the bar there is 5e-10
@pytest.mark.parametrize("size,N,bar", [((8, 5), 77, 1e-10), ((1, 1), 45, 1e-10), ((10, 11), 161, 5e-10)])
@pytest.mark.parametrize("path", ["stored", "fused"])
def test_odd_sizes_vs_oracle(oracle, size, N, bar, path):
    L = syn.length_unit()
    P = syn.thermal_splats(2, A, 5)
    if size == (1, 1):      # the brightest pixel of a coarse oracle image
        coarse = K.Camera((-9.0, 9.0), (-9.0, 9.0), 9)
        I = oracle.polarized_image(A, TH, coarse.alphas, coarse.betas, N, P, 0.0, 230e9, L)[0][:, 0]
        j = int(np.argmax(I))
        cam = K.Camera((coarse.alphas[j], coarse.alphas[j]), (coarse.betas[j], coarse.betas[j]), size)
    else:
        cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), size)
    c = cache_of(cam, N, path)
    img = K.polarized_image(c, P, 0.0, 230e9, L)
    ref = oracle.polarized_image(A, TH, cam.alphas, cam.betas, N, P, 0.0, 230e9, L)[0].reshape(cam.size + (4,), order="F")
    peak = np.abs(ref[..., 0]).max()
    assert peak > 0
    assert np.abs(img - ref).max() <= bar * peak
    if path != "stored":
        return
    ds = np.random.default_rng(3).normal(size=cam.size + (1, 1, 4))
    g, _ = K.polarized_gradient(c, P, [0.0], [230e9], L, ds)
    gref, _ = oracle.polarized_gradient(A, TH, cam.alphas, cam.betas, N, P, 0.0, 230e9, L, ds[:, :, 0, 0, :].reshape(cam.npix, 4, order="F"))
    assert np.abs(gref).max() > 0
    assert np.abs(g - gref).max() <= 1e-9 * np.abs(gref).max()
    K.precision(c, "mixed")
    g32, img32 = K.polarized_gradient(c, P, [0.0], [230e9], L, ds)
    assert np.abs(img32[:, :, 0, 0] - img).max() <= 1e-4 * peak
    assert np.abs(g32 - g).max() <= 1e-3 * np.abs(g).max()


def test_zero_angular_momentum_column(oracle):
    """A 7 x 5 screen (alpha = 0 in its middle column): every other pixel at the FP64 bar; the middle column's
    device values are finite (a NaN phi drops the ray's samples, leaving the pixel dark or partly lit)."""
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), (7, 5))
    L = syn.length_unit()
    P = syn.thermal_splats(2, A, 5)
    c = cache_of(cam, 200, "stored")
    img = K.polarized_image(c, P, 0.0, 230e9, L)
    ref = oracle.polarized_image(A, TH, cam.alphas, cam.betas, 200, P, 0.0, 230e9, L)[0].reshape(cam.size + (4,), order="F")
    peak = np.abs(ref[..., 0]).max()
    assert np.all(np.isfinite(img))
    rest = np.ones(cam.size, dtype=bool)
    rest[3, :] = False
    assert np.abs(img - ref)[rest].max() <= 1e-10 * peak
