"""The oracle's generic dual-sweep gradient (every row of every splat as a dual number through the same element
the transport evaluates) for every population, against central finite differences of the oracle's renderer.

Cases: power-law splats (22 rows: p, ln gamma_min), kappa splats (22 rows: kappa, ln w; the 2F1 factor is
recomputed from kappa and w at every evaluation, so its derivative is part of the gradient), kappa splats whose
centres follow knots (the centre rows translate the whole trajectory; the pattern rate does not enter), thermal
knot splats, and a composite of thermal, power-law and kappa parts. Dense parcels, so that absorption and
Faraday rotation act at first order. Step 1e-6 relative (5e-8 absolute near zero: the pattern rate, ~0.05, multiplies emission times of hundreds of M, so its curvature is large), every row differenced, the
bar 1e-6 of the largest gradient entry (the truncation and rounding noise of the differences is ~1e-9).
The thermal case is also checked against the seven-partial thermal path (the same gradient by another route).
"""
import math

import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn
from kerrsplat_amd.api import Camera

A, THO = 0.5, math.radians(60.0)
NU = 230e9
N = 128


def camera():
    return Camera((-15, 15), (-15, 15), 9)


def knots_for(P, times):
    """Knots on a straight drift of each splat's centre (any path works: the interpolation is what is tested)."""
    ns = P.shape[1]
    v = np.array([0.02, -0.015, 0.005])
    return np.stack([P[0:3, i][:, None] + np.outer(v, times) for i in range(ns)], axis=2)


def case_parts(case, oracle):
    if case == "powerlaw":
        P = syn.powerlaw_splats(1, A, 2)
        P[12, :] += math.log(300.0)
        return [("powerlaw", P, None, 1e5)]
    if case in ("kappa", "kappa_knots"):
        P = syn.kappa_splats(1, A, 2)
        P[12, :] += math.log(30.0)
        hyp = [oracle.kappa_hyp(P[13, i], math.exp(P[14, i])) for i in range(P.shape[1])]
        if case == "kappa":
            return [("kappa", P, hyp, 1e5)]
        times = np.array([-300.0, -100.0, 20.0, 150.0])
        return [("kappa", P, hyp, 1e5, knots_for(P, times), times)]
    if case == "thermal_knots":
        P = syn.thermal_splats(1, A, 2)
        P[12, :] += math.log(30.0)
        times = np.array([-250.0, -60.0, 40.0])
        return [("knot", P, knots_for(P, times), times)]
    if case == "composite":
        Pt = syn.thermal_splats(1, A, 1); Pt[12, :] += math.log(30.0)
        Pp = syn.powerlaw_splats(2, A, 1); Pp[12, :] += math.log(300.0)
        Pk = syn.kappa_splats(3, A, 1); Pk[12, :] += math.log(30.0)
        hyp = [oracle.kappa_hyp(Pk[13, 0], math.exp(Pk[14, 0]))]
        return [("thermal", Pt), ("powerlaw", Pp, None, 1e5), ("kappa", Pk, hyp, 1e5)]
    raise ValueError(case)


def with_params(parts, q, Q, oracle):
    """parts with part q's parameters replaced by Q (the kappa 2F1 factors recomputed from Q)."""
    out = list(parts)
    part = list(out[q])
    part[1] = Q
    if part[0] == "kappa":
        part[2] = [oracle.kappa_hyp(Q[13, i], math.exp(Q[14, i])) for i in range(Q.shape[1])]
    out[q] = tuple(part)
    return out


@pytest.mark.parametrize("case", ["powerlaw", "kappa", "kappa_knots", "thermal_knots", "composite"])
def test_generic_gradient_vs_central_differences(oracle, case):
    cam = camera()
    L = syn.length_unit()
    parts = case_parts(case, oracle)
    ds = np.random.default_rng(7).normal(size=(cam.npix, 4))
    grads, img = oracle.model_gradient(A, THO, cam.alphas, cam.betas, N, parts, 0.0, NU, L, ds)
    ref, _, _ = oracle.polarized_composite(A, THO, cam.alphas, cam.betas, N, parts, 0.0, NU, L)
    assert np.abs(img - ref).max() <= 1e-14 * np.abs(ref).max()

    def loss(pp):
        im, _, _ = oracle.polarized_composite(A, THO, cam.alphas, cam.betas, N, pp, 0.0, NU, L)
        return float(np.sum(ds * im))

    gmax = max(np.abs(g).max() for g in grads)
    assert gmax > 0
    worst = 0.0
    for q, part in enumerate(parts):
        P = np.asarray(part[1])
        for i in range(P.shape[1]):
            for r in range(P.shape[0]):
                h = 1e-6 * max(abs(P[r, i]), 0.05)
                Qp, Qm = P.copy(), P.copy()
                Qp[r, i] += h
                Qm[r, i] -= h
                if part[0] == "knot" or len(part) > 5:
                    if r < 3:           # the centre rows translate the knots rigidly
                        kp, km = np.array(part[-2], copy=True), np.array(part[-2], copy=True)
                        kp[r, :, i] += h
                        km[r, :, i] -= h
                        pp = with_params(parts, q, Qp, oracle); pm = with_params(parts, q, Qm, oracle)
                        pp[q] = pp[q][:-2] + (kp, part[-1]); pm[q] = pm[q][:-2] + (km, part[-1])
                        fd = (loss(pp) - loss(pm)) / (2 * h)
                        worst = max(worst, abs(fd - grads[q][r, i]) / gmax)
                        continue
                fd = (loss(with_params(parts, q, Qp, oracle)) - loss(with_params(parts, q, Qm, oracle))) / (2 * h)
                worst = max(worst, abs(fd - grads[q][r, i]) / gmax)
    assert worst < 1e-6, (case, worst)
    if case in ("kappa_knots", "thermal_knots"):
        # no pattern rotation for knot splats: the rate's row has no effect
        assert np.all(grads[0][-1] == 0.0)


def test_generic_gradient_equals_the_thermal_path(oracle):
    cam = camera()
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(30.0)
    ds = np.random.default_rng(3).normal(size=(cam.npix, 4))
    g7, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, P, 0.0, NU, L, ds)
    (gg,), _ = oracle.model_gradient(A, THO, cam.alphas, cam.betas, N, [("thermal", P)], 0.0, NU, L, ds)
    assert np.abs(gg - g7).max() <= 1e-12 * np.abs(g7).max()


def test_kappa_hypergeometric_against_scipy(oracle):
    """The kappa 2F1 factor (whose kappa and w derivatives the kappa rows carry) against scipy, value and w slope."""
    import scipy.special as sps
    for kap, w in [(3.2, 4.0), (5.5, 12.0), (6.9, 27.0)]:
        a, b, c, z = kap - 1 / 3, kap + 1, kap + 2 / 3, -kap * w
        assert abs(oracle.kappa_hyp(kap, w) - sps.hyp2f1(a, b, c, z)) <= 1e-12 * abs(sps.hyp2f1(a, b, c, z))
        hw = 1e-4 * w
        dw = (oracle.kappa_hyp(kap, w + hw) - oracle.kappa_hyp(kap, w - hw)) / (2 * hw)
        dw_exact = a * b / c * sps.hyp2f1(a + 1, b + 1, c + 1, z) * (-kap)
        assert abs(dw - dw_exact) <= 1e-6 * abs(dw_exact)
