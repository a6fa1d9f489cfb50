"""The full fit step's model (a = 0.9, theta_o = 163 deg, 1024 kappa splats whose centres follow 4 trajectory knots
from -100 to 200 M, 3 frequencies x 32 frames over 100 M) on the device.

* Reduced screen (6 x 6 over +-15 M, N = 256), one (frame, frequency) slice: the device gradient with respect to
  every row of every splat against the oracle's generic dual-sweep gradient (every row a dual number through the
  element; pinned by central differences in tests/test_oracle_adjoint_populations.py), FP64 at 1e-9 of the largest
  entry, mixed at 1e-3; the images at 1e-10 (FP64) and 1e-4 (mixed) of the peak.
* 128 x 128, N = 512, three slices (first, middle and last frame at the three frequencies): every image and
  gradient entry finite on both paths (the power-law and kappa fits diverge where the photon travels along the
  field; the pitch-angle sine is floored where it rounds to zero, and the sin^-p - 1 factors are taken from the
  cosine near sin = 1), and the mixed path within 1e-4 of the peak Stokes I and 1e-3 of the largest gradient entry
  of the FP64 path.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

C = syn.CONFIGS[4]
A, THO = C["a"], math.radians(C["theta_o_deg"])


def setup(res, N):
    cam = K.Camera((-C["fov"], C["fov"]), (-C["fov"], C["fov"]), res)
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    _, P, kn, kt = syn.config_model(4)
    return cam, cache, P, kn, kt


@pytest.fixture(scope="module")
def reduced(oracle):
    res, N, t, nu = 6, 256, 50.0, 230e9
    cam, cache, P, kn, kt = setup(res, N)
    L = syn.length_unit()
    ds = np.random.default_rng(41).normal(size=(res, res, 1, 1, 4))
    dso = ds[:, :, 0, 0, :].reshape(-1, 4, order="F")
    hyp = [oracle.kappa_hyp(P[13, i], math.exp(P[14, i])) for i in range(P.shape[1])]
    print(f"cfg4 reduced: {res}x{res} x {N}, 1024 kappa knot splats, oracle gradient", flush=True)
    ref, img = oracle.model_gradient(A, THO, cam.alphas, cam.betas, N, [("kappa", P, hyp, 1e5, kn, kt)], t, nu, L, dso)
    return cam, cache, P, kn, kt, L, t, nu, ds, ref[0], img.reshape(res, res, 4, order="F")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cfg4_reduced_gradient_vs_oracle(reduced, precision):
    cam, cache, P, kn, kt, L, t, nu, ds, ref, img = reduced
    m = K.Model(P, kind="kappa", knots=kn, knot_times=kt)
    K.precision(cache, precision)
    try:
        print(f"cfg4 reduced gradient, {precision}", flush=True)
        g, cube = K.polarized_gradient(cache, m, [t], [nu], L, ds)
    finally:
        K.precision(cache, "fp64")
    gmax = np.abs(ref).max()
    assert gmax > 0 and np.isfinite(g).all()
    err = np.abs(g - ref).max() / gmax
    assert err <= (1e-9 if precision == "fp64" else 1e-3), err
    peak = np.abs(img[..., 0]).max()
    ierr = np.abs(cube[:, :, 0, 0, :] - img).max() / peak
    assert ierr <= (1e-10 if precision == "fp64" else 1e-4), ierr


def test_cfg4_finite_and_mixed_within_bars():
    res, N = 128, 512
    cam, cache, P, kn, kt = setup(res, N)
    L = syn.length_unit()
    m = K.Model(P, kind="kappa", knots=kn, knot_times=kt)
    ds = np.random.default_rng(3).normal(size=(res, res, 1, 1, 4))
    slices = [(C["times"][0], C["nus"][0]), (C["times"][15], C["nus"][1]), (C["times"][31], C["nus"][2])]
    out = {}
    for prec in ("fp64", "mixed"):
        K.precision(cache, prec)
        for t, nu in slices:
            print(f"cfg4 {res}x{res} x {N}: {prec}, t = {t:.1f}, nu = {nu:.3g}", flush=True)
            g, cube = K.polarized_gradient(cache, m, [t], [nu], L, ds)
            assert np.isfinite(cube).all() and np.isfinite(g).all()
            out[(prec, t)] = (cube, g)
    K.precision(cache, "fp64")
    for t, _ in slices:
        c64, g64 = out[("fp64", t)]
        c32, g32 = out[("mixed", t)]
        peak = np.abs(c64[..., 0]).max()
        assert peak > 0
        assert np.abs(c32 - c64).max() <= 1e-4 * peak
        assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()
