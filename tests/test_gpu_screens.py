"""Batched screens and the transfer's launch choices on the device (a reduced form of the batched configuration:
a 32 x 32 screen over +-12 M in uneven 8^2-16^2 tiles plus two extra screens of 10^2 and 16^2, one cache over the
concatenated points, a = 0.9, theta_o = 45 deg, 72 splats of the three populations, 2 frames x 2 frequencies).

* The concatenated cache against the oracle's composite render of the same points (FP64 at 1e-10 of the peak,
  mixed at 1e-4), and every tile and extra screen against a cache of its own (1e-13 of the peak: the items are
  computed alike, only the case sort and the wavefront grouping differ).
* The wavefront patch and the launch form (one workgroup per task, or the persistent work queue in index order
  and then in the cost order of the previous launch) change nothing: bit-identical cubes. The samples per work
  item (16, 64) change only the order of the Stokes sums (1e-12 of the peak in FP64, 1e-5 in the mixed path).
* Every task of a work-queue launch records its ticks; the cost-ordered launch repeats the result.
* The dual-sweep gradient with a tuned launch (chunk 64, queue) equals the default one (1e-12 of the largest
  entry); the reverse pass takes the chunk length of its forward pass; the element adjoints' persistent form gives
  the gradient bit for bit.
* The autotuner times all 24 candidates and keeps the fastest; the tuned cube equals the default one.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

A, TH = 0.9, math.radians(45.0)
TIMES, NUS = [0.0, 30.0], [86e9, 345e9]
N = 192


def parts(oracle):
    T = syn.thermal_splats(1, A, 24, single_frame=False)
    P = syn.powerlaw_splats(2, A, 24, single_frame=False)
    Kp = syn.kappa_splats(3, A, 24, single_frame=False)
    hyp = [oracle.kappa_hyp(Kp[13, i], math.exp(Kp[14, i])) for i in range(Kp.shape[1])]
    model = K.CompositeModel(K.Model(T), K.Model(P, kind="powerlaw"), K.Model(Kp, kind="kappa"))
    return model, [("thermal", T), ("powerlaw", P), ("kappa", Kp, hyp)]


@pytest.fixture(scope="module")
def batch(oracle):
    main, splits, extras = syn.config_screens(2, res=32, extra=2, extra_size=(10, 16), tile_sizes=(8, 16))
    main = K.Camera((-12.0, 12.0), (-12.0, 12.0), 32)
    tiles, idx = K.screen_tiles(main, splits)
    cams = tiles + extras
    cam = K.concatenate_screens(*cams)
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, A, TH, marcher=K.Recurrence(64))
    model, ps = parts(oracle)
    L = syn.length_unit()
    print(f"screens: {len(tiles)} tiles + {len(extras)} screens, {cam.npix} points, N = {N}", flush=True)
    cube = K.polarized_cube(cache, model, TIMES, NUS, L)
    return dict(cams=cams, cam=cam, cache=cache, model=model, ps=ps, L=L, cube=cube, main=main, idx=idx, ntiles=len(tiles))


def test_concatenated_vs_oracle(batch, oracle):
    cam, cache, cube, L = batch["cam"], batch["cache"], batch["cube"], batch["L"]
    assert cube.shape == (cam.npix, 1, 2, 2, 4)
    # the oracle on the last frame at the higher frequency
    ref = oracle.polarized_composite(A, TH, cam.alphas, cam.betas, N, batch["ps"], TIMES[1], NUS[1], L)[0]
    got = cube[:, 0, 1, 1, :]
    peak = np.abs(ref[:, 0]).max()
    assert peak > 0
    assert np.abs(got - ref).max() <= 1e-10 * peak
    K.precision(cache, "mixed")
    try:
        c32 = K.polarized_cube(cache, batch["model"], TIMES, NUS, L)
    finally:
        K.precision(cache, "fp64")
    assert np.abs(c32[:, 0, 1, 1, :] - ref).max() <= 1e-4 * peak


def test_parts_match_their_own_caches(batch):
    cams, cube, L = batch["cams"], batch["cube"], batch["L"]
    peak = np.abs(cube[..., 0]).max()
    off = 0
    for q, c in enumerate(cams):
        own = K.GeodesicCache(c, N, store_samples=True)
        K.regenerate(own, A, TH, marcher=K.Recurrence(64))
        sub = K.polarized_cube(own, batch["model"], TIMES, NUS, L)
        n = c.npix
        part = cube[off:off + n, 0]
        assert np.abs(sub.reshape(n, 2, 2, 4, order="F") - part).max() <= 1e-13 * peak, q
        off += n
    # the tiles put back together are the full screen
    main = batch["main"]
    full = K.GeodesicCache(main, N, store_samples=True)
    K.regenerate(full, A, TH, marcher=K.Recurrence(64))
    ref = K.polarized_cube(full, batch["model"], TIMES, NUS, L).reshape(main.npix, 2, 2, 4, order="F")
    back = np.zeros_like(ref)
    off = 0
    for q in range(batch["ntiles"]):
        n = cams[q].npix
        back[batch["idx"][q]] = cube[off:off + n, 0]
        off += n
    assert np.abs(back - ref).max() <= 1e-13 * peak


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_patch_and_queue_bit_identical(batch, precision):
    cache, model, L = batch["cache"], batch["model"], batch["L"]
    K.precision(cache, precision)
    try:
        K.tuning(cache, chunk=32, queue=False, patch=(8, 8))
        base = K.polarized_cube(cache, model, TIMES, NUS, L)
        for patch in [(16, 4), (4, 16), (64, 1)]:
            for queue in (False, True):
                print(f"{precision}: patch {patch}, queue {queue}", flush=True)
                K.tuning(cache, patch=patch, queue=queue)
                got = K.polarized_cube(cache, model, TIMES, NUS, L)
                assert np.array_equal(got, base), (patch, queue)
                if queue:
                    cost = K.queue_costs(cache)
                    assert len(cost) > 0 and (cost > 0).all()
                    # the next launch takes the tasks in the cost order of this one
                    again = K.polarized_cube(cache, model, TIMES, NUS, L)
                    assert np.array_equal(again, base)
        # the chunk length changes the order of the Stokes sums: rounding of the path's scalar
        peak = np.abs(base[..., 0]).max()
        tol = 1e-12 if precision == "fp64" else 1e-5
        for chunk in (16, 64):
            K.tuning(cache, chunk=chunk, queue=True, patch=(8, 8))
            got = K.polarized_cube(cache, model, TIMES, NUS, L)
            print(f"{precision}: chunk {chunk}: {np.abs(got - base).max() / peak:.2e} of the peak", flush=True)
            assert np.abs(got - base).max() <= tol * peak, chunk
    finally:
        K.tuning(cache, chunk=32, queue=False, patch=(8, 8))
        K.precision(cache, "fp64")


def test_gradient_with_tuned_launch(batch):
    cache, model, L = batch["cache"], batch["model"], batch["L"]
    cam = batch["cam"]
    ds = np.random.default_rng(9).normal(size=(cam.npix, 1, 2, 2, 4))
    g0, c0 = K.polarized_gradient(cache, model, TIMES, NUS, L, ds)
    K.tuning(cache, chunk=64, queue=True, patch=(16, 4))
    try:
        g1, c1 = K.polarized_gradient(cache, model, TIMES, NUS, L, ds)
        # a second forward + reverse with the cost-ordered queue, then the element adjoints' persistent form
        g2, _ = K.polarized_gradient(cache, model, TIMES, NUS, L, ds)
        K.tuning(cache, adjoint_queue=True)
        g3, _ = K.polarized_gradient(cache, model, TIMES, NUS, L, ds)
    finally:
        K.tuning(cache, chunk=32, queue=False, patch=(8, 8), adjoint_queue=False)
    gmax = np.abs(g0).max()
    assert gmax > 0
    assert np.abs(g1 - g0).max() <= 1e-12 * gmax
    assert np.array_equal(g2, g1)
    assert np.array_equal(g3, g1)
    assert np.abs(c1 - c0).max() <= 1e-12 * np.abs(c0[..., 0]).max()


def test_autotune(batch):
    cache, model, L = batch["cache"], batch["model"], batch["L"]
    choice, table = K.autotune(cache, model, TIMES, NUS, L, records=False, sample_every=2)
    assert len(table) == 24
    assert all(ms > 0 and math.isfinite(ms) for *_, ms in table)
    best = min(table, key=lambda r: r[3])
    assert (choice["chunk"], choice["patch"], choice["queue"]) == best[:3]
    print(f"autotune: {choice}", flush=True)
    got = K.polarized_cube(cache, model, TIMES, NUS, L)
    peak = np.abs(batch["cube"][..., 0]).max()
    assert np.abs(got - batch["cube"]).max() <= 1e-12 * peak
    K.tuning(cache, chunk=32, queue=False, patch=(8, 8))
    with pytest.raises(ValueError):
        K.tuning(cache, chunk=4)
    with pytest.raises(ValueError):
        K.tuning(cache, patch=(0, 8))
