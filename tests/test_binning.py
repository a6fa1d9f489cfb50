"""Pixel integration (Binning, bin, binned_grid, binned_polar, concatenate: camera.jl:51-164), screen
tiles (tiles: cache.jl:234-246) and the binned χ² seed of _frame_seed! (Fit.jl:555-574), on the host.

Pinned by what the definitions fix: sub-cell centres average to the pixel centre (grid) and to the
cell's mid radius (polar), bin() equals the reference's loop restated literally, the tiles cover the
screen in order with Julia's round-half-even bounds, and the binned seed is the gradient of the binned
χ² with respect to every point's image (central differences), which also reduces to the unbinned seed
2 r / σ when each pixel holds one point. `frame_seed_binned` below is the checker the GPU test uses.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K


def frame_seed_binned(points, data, sigma, mask, binning):
    """_frame_seed! through a Binning (Fit.jl:555-574): points (npoints, 4) in screen order, data and
    sigma (npixels, 4), mask (npixels,). Returns (χ², seed per point (npoints, 4))."""
    acc = np.zeros((binning.npixels, 4))
    for m in range(binning.npoints):
        acc[binning.pixel[m]] += points[m]
    stokes = acc / binning.count[:, None]
    r = (stokes - data) / sigma
    chi = float(np.sum((r * r)[mask]))
    pseed = np.where(mask[:, None], 2 * r / sigma / binning.count[:, None], 0.0)
    return chi, pseed[binning.pixel]


def test_binned_grid_points():
    cam, b = K.binned_grid((-3.0, 5.0), (-2.0, 2.0), (4, 5))
    da, db = 8.0 / 4, 4.0 / 5
    i, j = np.arange(20) % 4, np.arange(20) // 4
    assert np.allclose(cam.alphas, -3.0 + (i + 0.5) * da, rtol=0, atol=1e-15)
    assert np.allclose(cam.betas, -2.0 + (j + 0.5) * db, rtol=0, atol=1e-15)
    assert np.array_equal(b.pixel, np.arange(20)) and b.size == (4, 5)
    cam3, b3 = K.binned_grid((-3.0, 5.0), (-2.0, 2.0), (4, 5), subsamples=3)
    assert cam3.npix == 180 and np.all(b3.count == 9)
    assert np.array_equal(b3.pixel, np.repeat(np.arange(20), 9))              # pixel by pixel, screen order
    assert np.allclose(K.bin(b3, cam3.alphas).T.ravel(), cam.alphas, atol=1e-14)
    assert np.allclose(K.bin(b3, cam3.betas).T.ravel(), cam.betas, atol=1e-14)
    with pytest.raises(ValueError):
        K.binned_grid((0, 1), (0, 1), 3, subsamples=0)


def test_binned_polar_points():
    edges = [2.0, 3.0, 5.0]
    cam, b = K.binned_polar(edges, 6, subsamples=(2, 3), psi0=0.1)
    assert b.size == (2, 6) and np.all(b.count == 6) and cam.npix == 72
    rho, psi = np.hypot(cam.alphas, cam.betas), np.arctan2(cam.betas, cam.alphas)
    ir, ip = b.pixel % 2, b.pixel // 2
    assert np.all((rho > np.array(edges)[ir]) & (rho < np.array(edges)[ir + 1]))
    assert np.allclose(K.bin(b, rho).T.ravel(), np.repeat([[2.5, 4.0]], 6, axis=0).ravel())     # mid radius
    mid = 0.1 + (np.arange(6) + 0.5) * 2 * math.pi / 6
    dpsi = np.angle(np.exp(1j * (psi - mid[ip])))
    assert np.abs(dpsi).max() < math.pi / 6 and np.abs(K.bin(b, dpsi)).max() < 1e-14
    with pytest.raises(ValueError):
        K.binned_polar([1.0], 4)


def test_bin_matches_the_reference_loop():
    rng = np.random.default_rng(5)
    b = K.Binning(rng.integers(0, 12, size=200), (3, 4))
    x = rng.normal(size=(200, 4))
    acc = np.zeros((12, 4))
    for m in range(200):
        acc[b.pixel[m]] += x[m]
    ref = (acc / b.count[:, None]).reshape(4, 3, 4).transpose(1, 0, 2)          # pixel i + 3 j -> [i, j]
    assert np.array_equal(K.bin(b, x), ref)
    cube = rng.normal(size=(200, 1, 2, 3, 4))                                    # polarized_cube of a point camera
    out = K.bin(b, cube)
    assert out.shape == (3, 4, 2, 3, 4)
    assert np.array_equal(out[:, :, 1, 2], K.bin(b, cube[:, 0, 1, 2]))
    with pytest.raises(K.api.DimensionMismatch):
        K.bin(b, x[:10])


def test_binning_errors():
    with pytest.raises(ValueError, match="outside"):
        K.Binning([0, 1, 4], (2, 2))
    with pytest.raises(ValueError, match="at least one point"):
        K.Binning([0, 1, 1, 3], (2, 2))


def test_concatenate_and_tiles():
    c1, b1 = K.binned_grid((-1, 1), (-1, 1), (2, 2), subsamples=2)
    c2, b2 = K.binned_polar([1, 2], 3)
    c, b = K.concatenate((c1, b1), (c2, b2))
    assert b.size == (7, 1) and c.npix == 19
    assert np.array_equal(b.pixel, np.concatenate([b1.pixel, b2.pixel + 4]))
    assert np.array_equal(c.alphas, np.concatenate([c1.alphas, c2.alphas]))
    cam = K.Camera((-1, 1), (-1, 1), 5)
    ts = K.tiles(cam, 3)
    assert [len(r) for _, r in ts] == [8, 9, 8]                                 # round(25/3)=8, round(50/3)=17
    assert [i for _, r in ts for i in r] == list(range(25))
    for t, r in ts:
        assert np.array_equal(t.alphas, cam.alphas[r.start:r.stop]) and np.array_equal(t.betas, cam.betas[r.start:r.stop])
    assert len(K.tiles(cam, 40)) == 25                                           # empty tiles are dropped


def test_binned_seed_is_the_gradient():
    rng = np.random.default_rng(8)
    b = K.Binning(np.concatenate([np.arange(6), rng.integers(0, 6, size=14)]), (2, 3))
    pts = rng.normal(size=(20, 4))
    data, sigma = rng.normal(size=(6, 4)), rng.uniform(0.5, 2.0, size=(6, 4))
    mask = np.array([True, True, False, True, True, True])
    chi, seed = frame_seed_binned(pts, data, sigma, mask, b)
    h = 1e-6
    for m, k in [(0, 0), (2, 1), (7, 3), (19, 2), (13, 0)]:
        p, q = pts.copy(), pts.copy()
        p[m, k] += h
        q[m, k] -= h
        fd = (frame_seed_binned(p, data, sigma, mask, b)[0] - frame_seed_binned(q, data, sigma, mask, b)[0]) / (2 * h)
        assert abs(fd - seed[m, k]) <= 1e-7 * max(1.0, abs(seed[m, k])), (m, k, fd, seed[m, k])
    assert np.all(seed[b.pixel == 2] == 0)                                      # masked pixel
    # one point per pixel: the unbinned seed 2 r / sigma
    b1 = K.Binning(np.arange(6), (2, 3))
    chi1, seed1 = frame_seed_binned(pts[:6], data, sigma, mask, b1)
    r = (pts[:6] - data) / sigma
    assert np.allclose(seed1, np.where(mask[:, None], 2 * r / sigma, 0.0), rtol=1e-15, atol=0)
