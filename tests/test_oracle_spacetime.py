"""The oracle's spacetime tangent (spacetime_valgrad / spacetime_jacobian, joint.jl:74-154): forward-mode
partials of every pixel's observed Stokes vector with respect to (a, theta_o, ln L), the duals carried through
the pixel constants, the march and the transport.

Pinned by what the mathematics fixes, as the reference pins its own (test_spacetime.jl:72-113, central
differences): the partials match central differences of the oracle's FP64 images in each variable
(Richardson-extrapolated from h = 2e-6 and 1e-6, both marchers; steps of 1e-5 can straddle a sample that
crosses the horizon guard or a splat's cutoff, a jump the duals rightly do not see), the value is the FP64
image, and in the optically thin limit the image is linear in the path length, so dI/d ln L = I. Comparing one variable's partial with another's differences (a mis-seeded dual)
fails the bar by orders of magnitude.
"""
import math

import numpy as np
import pytest

from kerrsplat_amd import synthetic as syn

A, TH, N = 0.9, math.radians(60.0), 64


@pytest.fixture(scope="module")
def pixels():
    fov, n = 9.0, 10
    al, be = np.meshgrid(np.linspace(-fov, fov, n), np.linspace(-fov, fov, n), indexing="ij")
    return al.ravel(order="F"), be.ravel(order="F")


def image(oracle, pixels, P, a, th, L, marcher, kind="thermal", hyp=None):
    return oracle.polarized_image(a, th, pixels[0], pixels[1], N, P, 0.0, 230e9, L, kind=kind, hyp=hyp, marcher=marcher)[0]


def central(oracle, pixels, P, marcher, h=2e-6, kind="thermal", hyp=None):
    """Central differences in each variable, Richardson-extrapolated from steps h and h / 2 (error O(h^4))."""
    L = syn.length_unit()

    def diff(k, hh):
        e = np.eye(3)[k] * hh
        ip = image(oracle, pixels, P, A + e[0], TH + e[1], L * math.exp(e[2]), marcher, kind, hyp)
        im = image(oracle, pixels, P, A - e[0], TH - e[1], L * math.exp(-e[2]), marcher, kind, hyp)
        return (ip - im) / (2 * hh)

    return np.stack([(4 * diff(k, h / 2) - diff(k, h)) / 3 for k in range(3)], axis=-1)


@pytest.mark.parametrize("marcher", [0, 1])
def test_tangent_matches_central_differences(oracle, pixels, marcher):
    P = syn.thermal_splats(1, A, 3)
    L = syn.length_unit()
    I, dI = oracle.spacetime_tangent(A, TH, pixels[0], pixels[1], N, P, 0.0, 230e9, L, marcher=marcher)
    ref = image(oracle, pixels, P, A, TH, L, marcher)
    assert np.abs(I - ref).max() <= 1e-13 * np.abs(ref[:, 0]).max()
    fd = central(oracle, pixels, P, marcher)
    for k in range(3):
        scale = np.abs(dI[:, :, k]).max()
        assert scale > 0
        assert np.abs(fd[:, :, k] - dI[:, :, k]).max() <= 5e-6 * scale, (k, np.abs(fd[:, :, k] - dI[:, :, k]).max() / scale)
    # a dual seeded on the wrong variable
    assert np.abs(fd[:, :, 1] - dI[:, :, 0]).max() > 1e-2 * np.abs(dI[:, :, 0]).max()


def test_tangent_powerlaw_matches_central_differences(oracle, pixels):
    P = syn.powerlaw_splats(2, A, 3)
    L = syn.length_unit()
    I, dI = oracle.spacetime_tangent(A, TH, pixels[0], pixels[1], N, P, 0.0, 230e9, L, kind="powerlaw", marcher=0)
    fd = central(oracle, pixels, P, 0, kind="powerlaw")
    for k in range(3):
        scale = np.abs(dI[:, :, k]).max()
        assert np.abs(fd[:, :, k] - dI[:, :, k]).max() <= 5e-6 * scale


def test_optically_thin_image_is_linear_in_L(oracle, pixels):
    """I(L) = L dI/dL + O(tau): the remainder of dI/d ln L - I falls with the density like the optical depth."""
    rem = []
    for scale in (1e-6, 1e-9):
        P = syn.thermal_splats(1, A, 3)
        P[12] += math.log(scale)
        I, dI = oracle.spacetime_tangent(A, TH, pixels[0], pixels[1], N, P, 0.0, 230e9, syn.length_unit(), marcher=1)
        peak = np.abs(I[:, 0]).max()
        assert peak > 0
        rem.append(np.abs(dI[:, :, 2] - I).max() / peak)
    assert rem[1] <= 5e-6, rem
    assert 1e2 <= rem[0] / rem[1] <= 1e4, rem
