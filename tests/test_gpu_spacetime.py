"""GPU parity of the spacetime tangent (spacetime_valgrad / spacetime_jacobian, joint.jl:74-154).

The device tangent by differences (spacetime_tangent_differences: FP64 cubes rendered on the device at the
spacetime and at steps h, h/2 either side in a, theta_o and ln L, Richardson-extrapolated) against the oracle's
forward-mode tangent on the same samples, Direct and Recurrence(64) (pinned by central differences and the thin-limit linearity in
tests/test_oracle_spacetime.py): values at the FP64 image bar (1e-10 of the peak), partials at 2e-5 of each
partial's largest entry (the oracle's own differences meet its duals at 5e-6). A composite's partials in the thin
limit are the sums of its parts'. Then the reference's own checks of the spacetime block (test_spacetime.jl:94-127):
the valgrad value equals the chi2 of a stored Recurrence(64) cache, its gradient matches central differences of the
device chi2, the Jacobian's residuals square to the same chi2 with 2 J^T r equal to the gradient, and the stored and
fused forms agree with and without the half-orbit cutoff (1e-10 of the residuals, 1e-8 of the Jacobian).

The device kernel of the dual tangent (three-partial duals through the pixel constants, the closed-form samples
and the transport) stays disabled: it took a GPU memory fault on its first launch, the cause is not found, and
ks_spacetime_tangent refuses to launch it; its test is kept, skipped.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

DISABLED = pytest.mark.skip(reason="the device tangent kernel took a GPU memory fault on its first launch; the cause is not "
                                   "found and ks_spacetime_tangent refuses to launch it")

A, TH = 0.9, math.radians(60.0)


def parts(oracle, kind, n):
    if kind == "thermal":
        return syn.thermal_splats(1, A, n), None
    if kind == "powerlaw":
        return syn.powerlaw_splats(2, A, n), None
    P = syn.kappa_splats(3, A, n)
    return P, [oracle.kappa_hyp(P[13, i], math.exp(P[14, i])) for i in range(n)]


@DISABLED
@pytest.mark.parametrize("kind,nmax", [("thermal", -1), ("thermal", 0), ("powerlaw", -1), ("kappa", -1)])
def test_tangent_vs_oracle(oracle, kind, nmax):
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), 10)
    N = 64
    L = syn.length_unit()
    cache = K.GeodesicCache(cam, N, store_samples=False)
    K.regenerate(cache, A, TH, marcher=K.Fused(64))
    P, hyp = parts(oracle, kind, 3)
    times, nus = [0.0, 15.0], [230e9, 345e9]
    cube, dcube = K.spacetime_tangent(cache, P, times, nus, L, nmax=nmax, slab=0.5, kind=kind)
    assert cube.shape == (10, 10, 2, 2, 4) and dcube.shape == (10, 10, 2, 2, 4, 3)
    for it, t in enumerate(times):
        for inu, nu in enumerate(nus):
            I, dI = oracle.spacetime_tangent(A, TH, cam.alphas, cam.betas, N, P, t, nu, L, kind=kind, hyp=hyp, nmax=nmax, slab=0.5,
                                             marcher=0)
            I = I.reshape(10, 10, 4, order="F")
            dI = dI.reshape(10, 10, 4, 3, order="F")
            peak = np.abs(I[..., 0]).max()
            assert peak > 0
            assert np.abs(cube[:, :, it, inu] - I).max() <= 1e-10 * peak
            for k in range(3):
                scale = np.abs(dI[..., k]).max()
                assert scale > 0
                err = np.abs(dcube[:, :, it, inu, :, k] - dI[..., k]).max() / scale
                assert err <= 1e-8, (kind, it, inu, k, err)


@pytest.mark.parametrize("kind,nmax,marcher", [("thermal", -1, 0), ("thermal", 0, 0), ("powerlaw", -1, 0), ("kappa", -1, 0),
                                                ("thermal", -1, 1), ("kappa", 0, 1)])
def test_tangent_differences_vs_oracle(oracle, kind, nmax, marcher):
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), 10)
    N = 64
    L = syn.length_unit()
    P, hyp = parts(oracle, kind, 3)
    times, nus = [0.0, 15.0], [230e9, 345e9]
    print(f"tangent by differences: {kind}, nmax {nmax}, marcher {marcher}", flush=True)
    march = K.Direct() if marcher == 0 else K.Recurrence(64)
    cube, dcube = K.spacetime_tangent_differences(cam, N, A, TH, P, times, nus, L, nmax=nmax, slab=0.5, kind=kind, marcher=march)
    assert cube.shape == (10, 10, 2, 2, 4) and dcube.shape == (10, 10, 2, 2, 4, 3)
    for it, t in enumerate(times):
        for inu, nu in enumerate(nus):
            I, dI = oracle.spacetime_tangent(A, TH, cam.alphas, cam.betas, N, P, t, nu, L, kind=kind, hyp=hyp, nmax=nmax, slab=0.5,
                                             marcher=marcher)
            I = I.reshape(10, 10, 4, order="F")
            dI = dI.reshape(10, 10, 4, 3, order="F")
            peak = np.abs(I[..., 0]).max()
            assert peak > 0
            assert np.abs(cube[:, :, it, inu] - I).max() <= 1e-10 * peak
            for k in range(3):
                scale = np.abs(dI[..., k]).max()
                assert scale > 0
                err = np.abs(dcube[:, :, it, inu, :, k] - dI[..., k]).max() / scale
                print(f"  t {t}, nu {nu:.3g}, partial {k}: {err:.2e} of its largest entry", flush=True)
                assert err <= 2e-5, (kind, it, inu, k, err)


def test_composite_tangent_is_the_sum_of_its_thin_parts(oracle):
    """A composite's partials in the optically thin limit are the sums of its parts' partials."""
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), 8)
    L = syn.length_unit()
    T, Pl = syn.thermal_splats(1, A, 2), syn.powerlaw_splats(2, A, 2)
    T[12] += math.log(1e-9)
    Pl[12] += math.log(1e-6)
    mt, mp = K.Model(T), K.Model(Pl, kind="powerlaw")
    c, dc = K.spacetime_tangent_differences(cam, 48, A, TH, K.CompositeModel(mt, mp), [0.0], [230e9], L)
    ct, dct = K.spacetime_tangent_differences(cam, 48, A, TH, mt, [0.0], [230e9], L)
    cp, dcp = K.spacetime_tangent_differences(cam, 48, A, TH, mp, [0.0], [230e9], L)
    for k in range(3):
        scale = np.abs(dc[..., k]).max()
        assert np.abs(dc[..., k] - dct[..., k] - dcp[..., k]).max() <= 1e-4 * scale
    assert np.abs(dct[..., 0]).max() > 1e-2 * np.abs(dc[..., 0]).max() and np.abs(dcp[..., 0]).max() > 1e-2 * np.abs(dc[..., 0]).max()


def test_spacetime_valgrad_and_jacobian():
    """test_spacetime.jl:94-113 on this build: value, gradient against central differences, Jacobian."""
    rng = np.random.default_rng(21)
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), 8)
    N = 40
    L = K.gravitational_radius(4e6)
    nu = 230e9
    p = syn.thermal_splats(1, 0.9, 2, single_frame=False)
    truth = K.GeodesicCache(cam, N, store_samples=False)
    K.regenerate(truth, 0.9, TH, marcher=K.Fused(64))
    times = [0.0, 10.0]
    clean = K.polarized_cube(truth, p, times, [nu], L)
    sigma = np.array([0.02, 0.01, 0.01, 0.005]) * np.abs(clean).max()
    data = clean + sigma * rng.normal(size=clean.shape)
    movie = K.StokesMovie(data, times, [nu], sigma)
    q = p.copy()
    q[12:15] += 0.05 * rng.normal(size=(3, 2))
    x3 = np.array([0.8, math.radians(52.0), math.log(1.2 * L)])
    v, g = K.spacetime_valgrad(K.spacetime_movie_loss(movie), x3, cam, N, q, times, [nu])
    stored = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(stored, x3[0], x3[1], marcher=K.Recurrence(64))
    chi_stored = K.chi2(q, movie, stored, math.exp(x3[2]))
    assert abs(v - chi_stored) <= 1e-8 * chi_stored

    def chi_at(y):
        c = K.GeodesicCache(cam, N, store_samples=True)
        K.regenerate(c, y[0], y[1], marcher=K.Direct())
        return K.chi2(q, movie, c, math.exp(y[2]))

    def diff(k, h):
        e = np.eye(3)[k] * h
        return (chi_at(x3 + e) - chi_at(x3 - e)) / (2 * h)

    h = 2e-6
    gfd = np.array([(4 * diff(k, h / 2) - diff(k, h)) / 3 for k in range(3)])
    e = np.max(np.abs(g - gfd) / np.maximum(np.abs(gfd), 1e-6 * np.abs(gfd).max()))
    assert e < 2e-5, (g, gfd, e)
    r, J = K.spacetime_jacobian(x3, cam, N, q, movie)
    assert abs(np.sum(r * r) - chi_stored) <= 1e-8 * chi_stored and J.shape == (len(r), 3)
    assert np.abs(2 * J.T @ r - g).max() <= 1e-8 * np.abs(g).max()
    # the stored and the fused forms agree, with and without the half-orbit truncation, and stay finite
    for nmax in (-1, 1):
        print(f"stored vs fused, nmax {nmax}", flush=True)
        rs, Js = K.spacetime_jacobian(x3, cam, N, q, movie, nmax=nmax, slab=0.5, stored=True)
        rf, Jf = K.spacetime_jacobian(x3, cam, N, q, movie, nmax=nmax, slab=0.5, stored=False)
        assert np.all(np.isfinite(Js)) and np.all(np.isfinite(Jf))
        print(f"  residuals {np.abs(rs - rf).max() / np.abs(rf).max():.2e}, Jacobian {np.abs(Js - Jf).max() / np.abs(Jf).max():.2e}", flush=True)
        assert np.abs(rs - rf).max() <= 1e-10 * np.abs(rf).max() and np.abs(Js - Jf).max() <= 1e-8 * np.abs(Jf).max()


def test_spacetime_errors():
    cam = K.Camera((-9.0, 9.0), (-9.0, 9.0), 4)
    cache = K.GeodesicCache(cam, 16, store_samples=False)
    with pytest.raises(K.api.StateError):
        K.spacetime_tangent(cache, syn.thermal_splats(1, A, 1), [0.0], [230e9], 1e14)
    K.regenerate(cache, A, TH, marcher=K.Fused(64))
    with pytest.raises(ValueError, match="thermal, power-law"):
        K.spacetime_tangent(cache, K.Model(syn.thin_splats(0, A, 1), kind="thin"), [0.0], [230e9], 1e14)
    with pytest.raises(ValueError, match="positive"):
        K.spacetime_tangent(cache, syn.thermal_splats(1, A, 1), [0.0], [230e9], -1.0)
    with pytest.raises(K.api.StateError, match="disabled"):                # refuses to launch the kernel
        K.spacetime_tangent(cache, syn.thermal_splats(1, A, 1), [0.0], [230e9], 1e14)
    with pytest.raises(ValueError, match="positive"):
        K.spacetime_tangent_differences(cam, 16, A, TH, syn.thermal_splats(1, A, 1), [0.0], [230e9], -1.0)
    with pytest.raises(ValueError, match="positive"):
        K.spacetime_tangent_differences(cam, 16, A, TH, syn.thermal_splats(1, A, 1), [0.0], [230e9], 1e14, h=0.0)
