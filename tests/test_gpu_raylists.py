"""GPU parity of the per-ray splat lists (ray_lists, dual_gradient.jl:283-359) against the oracle's hit table
(pinned in tests/test_oracle_raylists.py by brute force and by the transport over the lists): the same
splats for every ray and frame, in ascending order, -1 past each count, with and without the half-orbit
cutoff, thermal and power-law splats. Device and oracle march their samples separately (agreeing to about
1e-12), so (ray, splat) pairs closer than 1e-9 to the support sphere are left out of the comparison.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

from test_oracle_raylists import brute_force, splats

pytestmark = pytest.mark.gpu

A, TH, N = 0.5, math.radians(60.0), 96


@pytest.fixture(scope="module")
def setup():
    cam = K.Camera((-12.0, 12.0), (-12.0, 12.0), 10)
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, A, TH, marcher=K.Recurrence(64))
    return cam, cache


def as_hits(lists, ns):
    l = K.framed(lists)
    h = np.zeros((ns, l.ids.shape[1], l.nframes), dtype=np.uint8)
    for f in range(l.nframes):
        for j in range(l.ids.shape[1]):
            ids = l.ray(j, f)
            assert np.all(np.diff(ids) > 0) and np.all(l.ids[l.count[j, f]:, j, f] == -1)
            h[ids, j, f] = 1
    return h


@pytest.mark.parametrize("kind,nmax,slab", [("thermal", -1, 0.0), ("thermal", 0, 0.5), ("powerlaw", -1, 0.0)])
def test_lists_match_oracle(oracle, setup, kind, nmax, slab):
    cam, cache = setup
    P = splats()
    if kind == "powerlaw":
        Q = syn.powerlaw_splats(4, A, P.shape[1])
        Q[0:12], Q[21] = P[0:12], P[20]
        P = Q
    times = [0.0, 40.0]
    lists = K.ray_lists(cache, K.Model(P, kind=kind), times, nmax=nmax, slab=slab)
    assert lists.ids.shape[1:] == (cam.npix, 2) and lists.count.shape == (cam.npix, 2)
    assert lists.capacity == max(1, lists.count.max())
    got = as_hits(lists, P.shape[1])
    ref = oracle.ray_hits(A, TH, cam.alphas, cam.betas, N, P, times, nmax=nmax, slab=slab)
    assert np.array_equal(lists.count, got.sum(axis=0))
    keep = np.ones(ref.shape, dtype=bool)
    for f, t in enumerate(times):
        Pg = np.vstack([P[0:12], np.zeros((8, P.shape[1])), P[-1:]])      # geometry rows, pattern rate last
        _, margin = brute_force(oracle, (cam.alphas, cam.betas), Pg, t)
        keep[:, :, f] = np.abs(margin) > 1e-9
    assert np.array_equal(got[keep], ref[keep])
    assert keep.sum() >= keep.size - 4 and ref.sum() > 20


def test_scalar_time_and_errors(oracle, setup):
    cam, cache = setup
    P = splats()
    m = K.Model(P)
    one = K.ray_lists(cache, m, 40.0)
    two = K.ray_lists(cache, m, [0.0, 40.0])
    assert one.ids.ndim == 2 and one.count.shape == (cam.npix,)
    assert np.array_equal(one.count, two.count[:, 1])
    assert np.array_equal(as_hits(one, P.shape[1])[:, :, 0], as_hits(two, P.shape[1])[:, :, 1])
    f = K.frame_lists(two, 1)
    assert np.array_equal(f.count, one.count)
    fl = K.flattened(two)
    assert fl.count.shape == (2 * cam.npix,) and np.array_equal(fl.count[cam.npix:], one.count)
    fused = K.GeodesicCache(cam, N, store_samples=False)
    K.regenerate(fused, A, TH, marcher=K.Fused(64))
    with pytest.raises(K.api.StateError, match="stored samples"):
        K.ray_lists(fused, m, 0.0)
    with pytest.raises(ValueError, match="thermal, power-law or kappa"):
        K.ray_lists(cache, K.Model(syn.thin_splats(0, A, 3), kind="thin"), 0.0)
