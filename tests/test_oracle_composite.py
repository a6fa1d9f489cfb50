"""The oracle's CompositeModel (transport.jl:298-316), pinned by what the definition fixes.

* Concatenation: the elements of a composite are those of its parts in order, so a thermal set cut
  into two parts renders exactly the single set (the same sum in the same order), and an empty part
  changes nothing.
* Mixed populations: coefficients of overlapping elements add, so in the optically thin limit the
  composite's Stokes image is the sum of its parts' images, with a remainder of first order in the
  optical depth: scaling every density by 1e-3 and 1e-6 must shrink the remainder by about 1e3 per
  step. A part that is dropped, or rendered with the first part's population, breaks the sum by at
  least 1% of the peak here.
* Order: swapping parts permutes the sum, which leaves the image unchanged up to rounding.
"""
import math

import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn

A, TH, N, NU = 0.9, math.radians(45.0), 128, 230e9


@pytest.fixture(scope="module")
def pixels():
    fov, n = 12.0, 10
    al, be = np.meshgrid(np.linspace(-fov, fov, n), np.linspace(-fov, fov, n), indexing="ij")
    return al.ravel(order="F"), be.ravel(order="F")


def parts(oracle, dlogne=0.0):
    T, P, K = syn.thermal_splats(1, A, 4), syn.powerlaw_splats(2, A, 3), syn.kappa_splats(3, A, 3)
    for X in (T, P, K):
        X[12] += dlogne
    hyp = [oracle.kappa_hyp(K[13, i], math.exp(K[14, i])) for i in range(K.shape[1])]
    return [("thermal", T), ("powerlaw", P), ("kappa", K, hyp)]


def render(oracle, pixels, ps):
    return oracle.polarized_composite(A, TH, pixels[0], pixels[1], N, ps, 0.0, NU, syn.length_unit())[0]


def test_composite_split_is_the_single_set(oracle, pixels):
    T = syn.thermal_splats(1, A, 4)
    ref = oracle.polarized_image(A, TH, pixels[0], pixels[1], N, T, 0.0, NU, syn.length_unit())[0]
    assert ref[:, 0].max() > 0
    assert np.array_equal(render(oracle, pixels, [("thermal", T[:, :2]), ("thermal", T[:, 2:])]), ref)
    assert np.array_equal(render(oracle, pixels, [("thermal", T), ("thermal", T[:, :0])]), ref)
    assert np.array_equal(render(oracle, pixels, [("thermal", T[:, :0]), ("thermal", T)]), ref)


def test_composite_adds_populations_in_the_thin_limit(oracle, pixels):
    rem = []
    for dl in (math.log(1e-3), math.log(1e-6)):
        ps = parts(oracle, dl)
        comp = render(oracle, pixels, ps)
        each = [render(oracle, pixels, [p]) for p in ps]
        peak = comp[:, 0].max()
        assert peak > 0
        assert min(e[:, 0].max() for e in each) >= 1e-2 * peak       # every population carries flux
        rem.append(np.abs(comp - sum(each)).max() / peak)
        # a dropped part breaks the sum by its share of the flux
        for drop in range(3):
            kept = [p for q, p in enumerate(ps) if q != drop]
            assert np.abs(render(oracle, pixels, kept) - sum(each)).max() / peak > 1e-2
        # the power-law set rendered with the first part's population (thermal)
        wrong = render(oracle, pixels, [ps[0], ("thermal", ps[1][1]), ps[2]])
        assert np.abs(wrong - sum(each)).max() / peak > 1e-2
    assert rem[1] <= 5e-5, rem
    assert 1e2 <= rem[0] / rem[1] <= 1e4, rem                       # first order in the optical depth


def test_composite_order_is_rounding_only(oracle, pixels):
    ps = parts(oracle)
    comp = render(oracle, pixels, ps)
    assert np.abs(render(oracle, pixels, ps[::-1]) - comp).max() <= 1e-13 * comp[:, 0].max()


def knot_set(a, n, times, seed=4):
    """Thermal splats whose centres follow knots on their pattern circles at the given times."""
    P = syn.thermal_splats(seed, a, n)
    kn = np.zeros((3, len(times), n))
    for i in range(n):
        for k, t in enumerate(times):
            ph = P[20, i] * t
            kn[:, k, i] = [P[0, i] * math.cos(ph) - P[1, i] * math.sin(ph), P[0, i] * math.sin(ph) + P[1, i] * math.cos(ph), P[2, i]]
    return P, kn


def test_knot_parts_split_and_add(oracle, pixels):
    """Knot parts (KnotSplats, trajectories.jl:71-90) in a composite: a knot set cut in two renders the single
    set exactly, and a knot part beside a thermal part adds to it in the thin limit (first order)."""
    times = np.array([-200.0, -100.0, 0.0, 100.0])
    P, kn = knot_set(A, 4, times)
    L = syn.length_unit()
    ref = oracle.polarized_image(A, TH, pixels[0], pixels[1], N, P, 30.0, NU, L, kind="knot", knots=kn, knot_times=times)[0]
    assert ref[:, 0].max() > 0
    comp = oracle.polarized_composite(A, TH, pixels[0], pixels[1], N, [("knot", P[:, :2], kn[:, :, :2], times),
                                                                       ("knot", P[:, 2:], kn[:, :, 2:], times)], 30.0, NU, L)[0]
    assert np.array_equal(comp, ref)
    rem = []
    for scale in (1e-3, 1e-6):
        Q, T = P.copy(), syn.thermal_splats(1, A, 3)
        Q[12] += math.log(scale)
        T[12] += math.log(scale)
        parts_ = [("knot", Q, kn, times), ("thermal", T)]
        both = oracle.polarized_composite(A, TH, pixels[0], pixels[1], N, parts_, 30.0, NU, L)[0]
        each = [oracle.polarized_composite(A, TH, pixels[0], pixels[1], N, [p], 30.0, NU, L)[0] for p in parts_]
        peak = both[:, 0].max()
        assert min(e[:, 0].max() for e in each) >= 1e-2 * peak
        rem.append(np.abs(both - sum(each)).max() / peak)
    assert rem[1] <= 5e-5 and 1e2 <= rem[0] / rem[1] <= 1e4, rem
