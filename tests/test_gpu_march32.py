"""The FP32 bulk march (march_precision "mixed": the integrand rates of t~ and phi, their Simpson sums and the
closed-form pieces in FP32 on the well-conditioned segments, between FP64 closed-form anchors; FP64 Jacobi states and
positions) against the FP64 march.

* Samples of three spacetimes (a = 0.94 / 17 deg, 0.5 / 60 deg, 0.9 / 163 deg; 24 x 24 over +-12 M, N = 512): the
  positions and flags bit for bit (the Jacobi states stay FP64), t~ and phi of the valid samples within 1e-4 (their
  FP32 increments between FP64 anchors, summed in FP64), and some difference (the FP32 path runs).
* The full fit step's model (1024 kappa splats on 4 knots, a = 0.9, theta_o = 163 deg) at 128 x 128 x 512, three
  slices: the mixed path (FP32 march bulk, FP32 transfer and adjoint) against the plain FP64 path within 1e-4 of
  the peak Stokes I and 1e-3 of the largest gradient entry.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def marched(cam, N, a, th, prec):
    c = K.GeodesicCache(cam, N, store_samples=True)
    K.march_precision(c, prec)
    K.regenerate(c, a, th, marcher=K.Recurrence(64))
    return c


@pytest.mark.parametrize("a,thdeg", [(0.94, 17.0), (0.5, 60.0), (0.9, 163.0)])
def test_fp32_march_samples(a, thdeg):
    cam = K.Camera((-12.0, 12.0), (-12.0, 12.0), 24)
    th = math.radians(thdeg)
    s64 = K.samples(marched(cam, 512, a, th, "fp64"))
    s32 = K.samples(marched(cam, 512, a, th, "mixed"))
    # positions and flags come from the FP64 Jacobi states in both: bit for bit
    assert np.array_equal(s64["flags"], s32["flags"])
    assert np.array_equal(s64["r"], s32["r"]) and np.array_equal(s64["theta"], s32["theta"])
    ok = (s64["flags"] & 1) != 0
    assert ok.sum() > 0.2 * ok.size
    terr = np.abs(s32["t"][ok] - s64["t"][ok])
    perr = np.abs(s32["phi"][ok] - s64["phi"][ok])
    print(f"a = {a}, theta_o = {thdeg}: t {terr.max():.1e}, phi {perr.max():.1e}", flush=True)
    assert 0 < terr.max() <= 1e-4 and perr.max() <= 1e-4


def test_fp32_march_full_step_model_within_bars():
    C = syn.CONFIGS[4]
    a, th = C["a"], math.radians(C["theta_o_deg"])
    cam = K.Camera((-C["fov"], C["fov"]), (-C["fov"], C["fov"]), 128)
    _, P, kn, kt = syn.config_model(4)
    L = syn.length_unit()
    m = K.Model(P, kind="kappa", knots=kn, knot_times=kt)
    ds = np.random.default_rng(3).normal(size=(128, 128, 1, 1, 4))
    c64 = marched(cam, 512, a, th, "fp64")
    c32 = marched(cam, 512, a, th, "mixed")
    K.precision(c32, "mixed")
    for t, nu in [(C["times"][0], C["nus"][0]), (C["times"][15], C["nus"][1]), (C["times"][31], C["nus"][2])]:
        g64, i64 = K.polarized_gradient(c64, m, [t], [nu], L, ds)
        g32, i32 = K.polarized_gradient(c32, m, [t], [nu], L, ds)
        peak = np.abs(i64[..., 0]).max()
        ie, ge = np.abs(i32 - i64).max() / peak, np.abs(g32 - g64).max() / np.abs(g64).max()
        print(f"t = {t:.1f}, nu = {nu:.3g}: image {ie:.1e}, gradient {ge:.1e}", flush=True)
        assert np.isfinite(i32).all() and np.isfinite(g32).all()
        assert ie <= 1e-4 and ge <= 1e-3
