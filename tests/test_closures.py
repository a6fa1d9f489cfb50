"""Closure quantities and station gains (visibilities.jl:141-214), host side.

Pinned by what defines them: closure phases are invariant under station-based phase gains and log
closure amplitudes under amplitude gains (the reason they are used); a planted mistake (a leg not
conjugated) breaks the invariance. The closure χ² derivative with respect to the model visibilities
(the input of the device seed) is checked against central differences in Re V and Im V.
"""
import itertools

import numpy as np

from kerrsplat_amd import fit


def array_layout(nst, rng):
    """All baselines i < j of nst stations (1-based), random visibilities, the closing triangles and quadrangles."""
    pairs = list(itertools.combinations(range(1, nst + 1), 2))
    index = {p: n + 1 for n, p in enumerate(pairs)}

    def leg(a, b):
        return index[(a, b)] if a < b else -index[(b, a)]

    tri = [(leg(i, j), leg(j, k), leg(k, i)) for i, j, k in itertools.combinations(range(1, nst + 1), 3)]
    quad = [(leg(i, j), leg(k, l), leg(i, k), leg(j, l)) for i, j, k, l in itertools.combinations(range(1, nst + 1), 4)]
    V = rng.normal(size=(len(pairs), 4)) + 1j * rng.normal(size=(len(pairs), 4))
    s1 = np.array([p[0] for p in pairs]); s2 = np.array([p[1] for p in pairs])
    return V, tri, quad, s1, s2


def test_closures_are_gain_invariant():
    rng = np.random.default_rng(1)
    V, tri, quad, s1, s2 = array_layout(6, rng)
    gains = np.vstack([rng.normal(scale=0.3, size=6), rng.uniform(-np.pi, np.pi, size=6)])
    Vg = fit.apply_gains(V, gains, s1, s2)
    assert np.abs(Vg - V).max() > 0.1
    cp, cpg = fit.closure_phases(V, tri), fit.closure_phases(Vg, tri)
    assert np.abs(np.remainder(cpg - cp + np.pi, 2 * np.pi) - np.pi).max() < 1e-12
    la, lag = fit.log_closure_amplitudes(V, quad), fit.log_closure_amplitudes(Vg, quad)
    assert np.abs(lag - la).max() < 1e-12
    # planted mistake: the third leg of every triangle not conjugated breaks the invariance
    bad = [(a, b, abs(c)) for a, b, c in tri]
    d = np.remainder(fit.closure_phases(Vg, bad) - fit.closure_phases(V, bad) + np.pi, 2 * np.pi) - np.pi
    assert np.abs(d).max() > 0.1


def test_closure_chi2_gradient_vs_central_differences():
    rng = np.random.default_rng(2)
    V, tri, quad, s1, s2 = array_layout(5, rng)
    data = fit.ClosureData(np.zeros(len(V)), np.zeros(len(V)), tri, fit.closure_phases(V, tri) + rng.normal(scale=0.3, size=len(tri)),
                           rng.uniform(0.1, 0.3, len(tri)), quad, fit.log_closure_amplitudes(V, quad) + rng.normal(scale=0.2, size=len(quad)),
                           rng.uniform(0.1, 0.3, len(quad)))
    chi, w = fit.closure_chi2(V, data, grad=True)
    assert chi > 0 and np.abs(w[:, 1:]).max() == 0.0
    h = 1e-6
    worst = 0.0
    for k in range(len(V)):
        for part, unit in (("re", 1.0), ("im", 1j)):
            Vp, Vm = V.copy(), V.copy()
            Vp[k, 0] += h * unit
            Vm[k, 0] -= h * unit
            fd = (fit.closure_chi2(Vp, data) - fit.closure_chi2(Vm, data)) / (2 * h)
            an = w[k, 0].real if part == "re" else w[k, 0].imag
            worst = max(worst, abs(fd - an))
    assert worst < 1e-6 * np.abs(w).max(), worst
