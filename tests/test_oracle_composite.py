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
