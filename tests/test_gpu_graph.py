"""Fit iterations captured as one hipGraph (FitGraph) against the same iteration run launch by launch
(gradient_forward, each slice's chi2 seed, gradient_reverse) on the same parameters: chi2 and every gradient entry
bit for bit, replay after replay with new parameters and with eager work between the replays; captured again for a
model of another splat count (a hygiene pass) and after the cache's launch choices changed; the image loss (thermal
splats, FP64 and mixed) and the visibility loss (kappa splats)."""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

A, TH = 0.9, math.radians(45.0)
TIMES, NUS = [0.0, 20.0], [230e9, 345e9]


@pytest.fixture(scope="module")
def cache():
    cam = K.Camera((-12.0, 12.0), (-12.0, 12.0), 14)
    c = K.GeodesicCache(cam, 160, store_samples=True)
    K.regenerate(c, A, TH, marcher=K.Recurrence(64))
    return c


def eager_image(cache, model, data, sigma, L):
    na, nb = cache.screen_size
    K.gradient_forward(cache, model, TIMES, NUS, L, copy=False)
    chi = 0.0
    for inu in range(len(NUS)):
        for it in range(len(TIMES)):
            chi += K.image_chi2_seed(cache, it + len(TIMES) * inu, data[:, :, it, inu, :], sigma)
    g = np.zeros(model.grad_shape, order="F")
    K.gradient_reverse(cache, model, g)
    return chi, g


def perturbed(P, rng, scale=0.02):
    Q = P.copy()
    Q[4:7] += scale * rng.normal(size=Q[4:7].shape)        # the log widths
    Q[11:13] += scale * rng.normal(size=Q[11:13].shape)    # log n_e, log Theta_e
    return Q


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_image_loss_graph_matches_eager(cache, precision):
    L = syn.length_unit()
    P0 = syn.thermal_splats(1, A, 10, single_frame=False)
    na, nb = cache.screen_size
    K.precision(cache, precision)
    try:
        truth = K.polarized_cube(cache, K.Model(P0), TIMES, NUS, L)
        rng = np.random.default_rng(4)
        data = truth * (1 + 0.05 * rng.normal(size=truth.shape))
        sigma = [0.05 * np.abs(truth[..., 0]).max()] * 4
        graph = K.FitGraph(cache, TIMES, NUS, L, image=(data, sigma))
        m = K.Model(P0)
        P = P0
        for k in range(3):
            P = perturbed(P0, rng) if k else P0
            print(f"graph step {k} ({precision})", flush=True)
            chi_g, g_g = graph.step(m, P)
            chi_e, g_e = eager_image(cache, K.Model(P), data, sigma, L)      # eager work between the replays
            assert chi_g == chi_e and np.array_equal(g_g, g_e), (k, chi_g, chi_e, np.abs(g_g - g_e).max())
        assert graph.info() == {"captures": 1, "replays": 3}
        # a hygiene pass: fewer splats, a new model -> captured again
        m2 = K.Model(P[:, :7])
        chi_g, g_g = graph.step(m2, P[:, :7])
        chi_e, g_e = eager_image(cache, K.Model(P[:, :7]), data, sigma, L)
        assert chi_g == chi_e and np.array_equal(g_g, g_e)
        # new launch choices of the cache -> captured again; the work queue gives the same numbers
        K.tuning(cache, queue=True, chunk=64)
        chi_q, g_q = graph.step(m2, P[:, :7])
        assert graph.info()["captures"] == 3
        chi_e, g_e = eager_image(cache, K.Model(P[:, :7]), data, sigma, L)
        assert chi_q == chi_e and np.array_equal(g_q, g_e)
        chi_q2, g_q2 = graph.step(m2, P[:, :7])
        assert chi_q2 == chi_q and np.array_equal(g_q2, g_q)
    finally:
        K.tuning(cache, queue=False, chunk=32)
        K.precision(cache, "fp64")


def test_vis_loss_graph_matches_eager(cache):
    L = syn.length_unit()
    Pk = syn.kappa_splats(3, A, 8, single_frame=False)
    na, nb = cache.screen_size
    u, v = syn.uv_disc(3, 64)
    psize = syn.pixel_size_rad(-12.0, 12.0, na)
    m = K.Model(Pk, kind="kappa")
    K.gradient_forward(cache, m, TIMES, NUS, L, copy=False)
    V = K.visibilities(cache, len(TIMES) * len(NUS), psize, u, v)          # (frames, nuv, 4)
    rng = np.random.default_rng(8)
    D = np.zeros((len(TIMES), len(NUS), len(u), 4), dtype=complex)
    for inu in range(len(NUS)):
        for it in range(len(TIMES)):
            f = it + len(TIMES) * inu
            D[it, inu] = V[f] * (1 + 0.05 * rng.normal(size=V[f].shape))
    S = np.full(D.shape, 0.02 * np.abs(V[..., 0]).max())
    graph = K.FitGraph(cache, TIMES, NUS, L, vis=(psize, u, v, D, S))
    for k in range(2):
        Q = perturbed(Pk, rng) if k else Pk
        print(f"vis graph step {k}", flush=True)
        chi_g, g_g = graph.step(m, Q)
        me = K.Model(Q, kind="kappa")
        K.gradient_forward(cache, me, TIMES, NUS, L, copy=False)
        chi_e = 0.0
        for inu in range(len(NUS)):
            for it in range(len(TIMES)):
                chi_e += K.vis_chi2_seed(cache, it + len(TIMES) * inu, psize, u, v, D[it, inu], S[it, inu])
        g_e = np.zeros(me.grad_shape, order="F")
        K.gradient_reverse(cache, me, g_e)
        assert chi_g == chi_e and np.array_equal(g_g, g_e), (k, chi_g, chi_e)


def test_graph_errors(cache):
    L = syn.length_unit()
    with pytest.raises(ValueError):
        K.FitGraph(cache, TIMES, NUS, L)
    na, nb = cache.screen_size
    g = K.FitGraph(cache, [0.0], [230e9], L, image=(np.zeros((na, nb, 1, 1, 4)), [1.0] * 4))
    th = K.Model(syn.thermal_splats(1, A, 2))
    comp = K.CompositeModel(th, K.Model(syn.kappa_splats(3, A, 2), kind="kappa"))
    with pytest.raises(ValueError, match="one thermal"):
        g.step(comp)
    with pytest.raises(K.api.DimensionMismatch):
        g.step(th, np.zeros((21, 3)))
