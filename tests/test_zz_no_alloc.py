"""The visibility transform's entry point that allocates no device memory (ks_visibilities_noalloc) against
ks_visibilities: the same kernels over the same inputs, so the results must agree bit for bit, on two calls in a
row (the second smaller, into the fixed workspace the first left behind) and after an ordinary call; a workload
beyond the fixed workspace fails with DimensionMismatch before anything runs."""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

A, THO, NU = 0.5, math.radians(60.0), 230e9


def test_visibilities_noalloc_matches_allocating_entry_point():
    cam = K.Camera((-15.0, 15.0), (-12.0, 12.0), (24, 20))
    cache = K.GeodesicCache(cam, 128, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 4)
    K.polarized_cube(cache, P, [0.0, 40.0], [NU], L)
    psize = syn.pixel_size_rad(-15.0, 15.0, 24)
    for frames, nuv, seed in [(2, 700, 3), (1, 90, 4)]:
        u, v = syn.uv_disc(seed, nuv)
        tap = np.random.default_rng(seed).uniform(0.5, 1.0, nuv)
        Vn = K.visibilities(cache, frames, psize, u, v, taper=tap, noalloc=True)
        Va = K.visibilities(cache, frames, psize, u, v, taper=tap)
        print(f"noalloc: {frames} frames x {nuv} uv: max |V| = {np.abs(Va).max():.3e}", flush=True)
        assert np.abs(Va).max() > 0
        np.testing.assert_array_equal(Vn, Va)
        Vn2 = K.visibilities(cache, frames, psize, u, v, noalloc=True)
        np.testing.assert_array_equal(Vn2, K.visibilities(cache, frames, psize, u, v))
    u, v = syn.uv_disc(5, 25000)
    with pytest.raises(K.api.DimensionMismatch):
        K.visibilities(cache, 2, psize, u, v, noalloc=True)
