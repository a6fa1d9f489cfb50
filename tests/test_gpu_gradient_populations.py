"""GPU parity of the dual-sweep gradient for every population: power-law, kappa, kappa with knots, thermal knot
splats and a composite of three populations, against the oracle's generic gradient (every splat row as a dual
number through the element the transport evaluates; pinned by central differences in
tests/test_oracle_adjoint_populations.py). Bars: FP64 path 1e-9 of the largest gradient entry (the thermal
bar), the mixed path 1e-3; the images of the same forward pass 1e-10 of each bright pixel's I. A 72-splat kappa
knot set also runs through the transfer's per-tile cull (64 splats and more), whose knot spheres must keep every
splat that contributes.
"""
import math
import os
import sys

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

sys.path.insert(0, os.path.dirname(__file__))
from test_oracle_adjoint_populations import case_parts, knots_for  # noqa: E402

pytestmark = pytest.mark.gpu

A, THO, NU, N = 0.5, math.radians(60.0), 230e9, 128


def stored(res):
    cam = K.Camera((-15, 15), (-15, 15), res)
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    return cam, cache


def device_model(part):
    kind, P = part[0], np.asarray(part[1])
    if kind == "knot":
        return K.Model(P, kind="knot", knots=part[2], knot_times=part[3])
    if len(part) > 5:
        return K.Model(P, kind=kind, knots=part[4], knot_times=part[5])
    return K.Model(P, kind=kind)


def screen_seed(res, seed):
    ds = np.random.default_rng(seed).normal(size=(res, res, 1, 1, 4))
    return ds, ds[:, :, 0, 0, :].reshape(-1, 4, order="F")


@pytest.mark.parametrize("case", ["powerlaw", "kappa", "kappa_knots", "thermal_knots", "composite"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_population_gradient_vs_oracle(oracle, case, precision):
    res = 10
    cam, cache = stored(res)
    if precision == "mixed":
        K.precision(cache, "mixed")
    L = syn.length_unit()
    parts = case_parts(case, oracle)
    models = [device_model(p) for p in parts]
    m = models[0] if len(models) == 1 else K.CompositeModel(*models)
    ds, ds_flat = screen_seed(res, 11)
    print(f"population gradient: {case}, {precision}, {res}x{res} x {N}", flush=True)
    g, cube = K.polarized_gradient(cache, m, [0.0], [NU], L, ds)
    ref, img = oracle.model_gradient(A, THO, cam.alphas, cam.betas, N, parts, 0.0, NU, L, ds_flat)
    blocks = m.split_gradient(g) if isinstance(m, K.CompositeModel) else [g]
    gmax = max(np.abs(r).max() for r in ref)
    assert gmax > 0
    err = max(np.abs(b - r).max() for b, r in zip(blocks, ref)) / gmax
    bar = 1e-9 if precision == "fp64" else 1e-3
    assert err <= bar, err
    img = img.reshape(res, res, 4, order="F")
    I = np.abs(img[..., 0])
    bright = I > 1e-6 * I.max()
    ierr = (np.abs(cube[:, :, 0, 0, :] - img).max(axis=-1) / np.maximum(I, 1e-300))[bright].max()
    assert ierr < (1e-10 if precision == "fp64" else 1e-3), ierr


def test_kappa_knots_through_the_per_tile_cull(oracle):
    """72 kappa splats whose centres follow knots (the per-tile cull runs from 64 splats on and bounds each knot
    splat by the sphere holding its knots): the image against the oracle's at 1e-10 of each bright pixel's I."""
    res = 16
    cam = K.Camera((-15, 15), (-15, 15), res)
    cache = K.GeodesicCache(cam, 256, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    L = syn.length_unit()
    P = syn.kappa_splats(4, A, 72, single_frame=False)
    times = np.array([-400.0, -150.0, 0.0, 120.0])
    kn = knots_for(P, times)
    m = K.Model(P, kind="kappa", knots=kn, knot_times=times)
    img = K.polarized_image(cache, m, 30.0, NU, L)
    hyp = [oracle.kappa_hyp(P[13, i], math.exp(P[14, i])) for i in range(P.shape[1])]
    ref, _, _ = oracle.polarized_image(A, THO, cam.alphas, cam.betas, 256, P, 30.0, NU, L, kind="kappa", hyp=hyp, knots=kn,
                                       knot_times=times)
    ref = ref.reshape(res, res, 4, order="F")
    I = np.abs(ref[..., 0])
    assert I.max() > 0
    bright = I > 1e-6 * I.max()
    err = (np.abs(img - ref).max(axis=-1) / np.maximum(I, 1e-300))[bright].max()
    assert err < 1e-10, err
