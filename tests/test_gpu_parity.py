"""GPU parity: the HIP library through its C-ABI against the CPU oracle on identical seeded inputs.

Bars: integer and index work (root cases, the case sort, sample flags) bit-exact; floating point
within the FP64 tolerances written in each test (per-sample geodesic quantities 1e-10 scaled by
max(|x|, 1); images 1e-10 of each pixel's Stokes I, the north-star bar for the plain FP64 path).
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def scaled_err(x, y):
    return np.abs(x - y) / np.maximum(np.abs(y), 1.0)


def camera(fov, res):
    return K.Camera((-fov, fov), (-fov, fov), res)


@pytest.mark.parametrize("a,thdeg,fov,res", [(0.94, 17.0, 12.0, 24), (0.5, 60.0, 15.0, 24), (0.9, 163.0, 10.0, 20)])
def test_case_sort_bit_exact(oracle, a, thdeg, fov, res):
    cam = camera(fov, res)
    th = math.radians(thdeg)
    cache = K.GeodesicCache(cam, 8, store_samples=False)
    K.regenerate(cache, a, th, marcher=K.Fused(64))
    nr = oracle.root_cases(a, th, cam.alphas, cam.betas)
    perm, counts = oracle.case_permutation(nr)
    assert tuple(counts) == cache.ranges
    assert np.array_equal(perm, cache.perm_host)


@pytest.mark.parametrize("marcher", ["direct", "recurrence"])
def test_samples_vs_oracle(oracle, marcher):
    a, th, N = 0.94, math.radians(17.0), 192
    cam = camera(12.0, 16)
    m = K.Direct() if marcher == "direct" else K.Recurrence(64)
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, a, th, marcher=m)
    got = K.samples(cache)
    idx = cache.perm_host - 1
    ref = oracle.samples(a, th, cam.alphas[idx], cam.betas[idx], N, marcher=0 if marcher == "direct" else 1)
    assert np.array_equal(got["flags"], ref["flags"])
    ok = (ref["flags"] & 1) == 1
    for f in ("r", "theta", "t", "phi"):
        e = scaled_err(got[f][ok], ref[f][ok])
        assert e.max() < 1e-10, (f, e.max())


def make_cache(cam, N, a, th, path):
    """A regenerated cache: 'stored' marches Recurrence(64) into sample storage (the transfer reads the
    samples), 'fused' marches inside every render (Fused(64), no storage)."""
    if path == "stored":
        cache = K.GeodesicCache(cam, N, store_samples=True)
        K.regenerate(cache, a, th, marcher=K.Recurrence(64))
    else:
        cache = K.GeodesicCache(cam, N, store_samples=False)
        K.regenerate(cache, a, th, marcher=K.Fused(64))
    return cache


PATHS = ["stored", "fused"]


@pytest.mark.parametrize("path", PATHS)
def test_thin_image_vs_oracle(oracle, path):
    cfg = syn.CONFIGS[0]
    a, th = cfg["a"], math.radians(cfg["theta_o_deg"])
    cam = camera(cfg["fov"], 24)
    N = 256
    P = syn.thin_splats(0, a, cfg["ns"])
    cache = make_cache(cam, N, a, th, path)
    img = K.thin_image(cache, P, 0.0)
    ref = oracle.thin_image(a, th, cam.alphas, cam.betas, N, P, 0.0).reshape(cam.size, order="F")
    assert np.max(np.abs(img - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.max(ref) > 0


def _polarized_case(oracle, kind, params, path, res=16, N=256, a=0.5, thdeg=60.0, fov=15.0, nu=230e9, nmax=-1, slab=0.0, **kw):
    th = math.radians(thdeg)
    cam = camera(fov, res)
    L = syn.length_unit()
    cache = make_cache(cam, N, a, th, path)
    model = K.Model(params, kind=kind, **kw)
    img = K.polarized_image(cache, model, 0.0, nu, L, nmax=nmax, slab=slab, kind=kind)
    hyp = None
    if kind == "kappa":
        hyp = [oracle.kappa_hyp(params[13, i], math.exp(params[14, i])) for i in range(params.shape[1])]
    ref, _, _ = oracle.polarized_image(a, th, cam.alphas, cam.betas, N, params, 0.0, nu, L, kind=kind, hyp=hyp,
                                       nmax=nmax, slab=slab, **kw)
    ref = ref.reshape(cam.size + (4,), order="F")
    I = np.abs(ref[..., 0])
    assert ref[..., 0].max() > 0
    err = np.abs(img - ref).max(axis=-1) / np.maximum(I, 1e-300)
    bright = I > 1e-6 * I.max()
    assert err[bright].max() < 1e-10, err[bright].max()
    assert np.abs(img - ref).max() <= 1e-10 * I.max()
    return img


@pytest.mark.parametrize("path", PATHS)
def test_polarized_thermal_vs_oracle(oracle, path):
    P = syn.thermal_splats(1, 0.5, 6)
    _polarized_case(oracle, "thermal", P, path)


@pytest.mark.parametrize("path", PATHS)
def test_polarized_thermal_winding_vs_oracle(oracle, path):
    P = syn.thermal_splats(1, 0.5, 6)
    _polarized_case(oracle, "thermal", P, path, nmax=0, slab=0.5)


@pytest.mark.parametrize("path", PATHS)
def test_polarized_powerlaw_vs_oracle(oracle, path):
    P = syn.powerlaw_splats(2, 0.9, 4)
    _polarized_case(oracle, "powerlaw", P, path, a=0.9, thdeg=45.0, fov=12.0)


@pytest.mark.parametrize("path", PATHS)
def test_polarized_kappa_vs_oracle(oracle, path):
    P = syn.kappa_splats(2, 0.9, 4)
    _polarized_case(oracle, "kappa", P, path, a=0.9, thdeg=45.0, fov=12.0)


@pytest.mark.parametrize("path", PATHS)
def test_polarized_knot_vs_oracle(oracle, path):
    P = syn.thermal_splats(4, 0.9, 3)
    times = np.array([-200.0, -100.0, 0.0, 100.0])
    knots = np.zeros((3, len(times), P.shape[1]))
    for i in range(P.shape[1]):
        for k, t in enumerate(times):
            ph = P[20, i] * t
            knots[:, k, i] = [P[0, i] * math.cos(ph) - P[1, i] * math.sin(ph), P[0, i] * math.sin(ph) + P[1, i] * math.cos(ph), P[2, i]]
    _polarized_case(oracle, "knot", P, path, a=0.9, thdeg=163.0, fov=12.0, knots=knots, knot_times=times)


def test_polarized_many_splats_tile_cull_vs_oracle(oracle):
    """72 thermal splats: the transfer over stored samples takes the per-tile cull (candidate lists from 64
    splats on) and the element adjoints the per-unit cull; both must leave the images and the gradient
    at the FP64 bars, and the mixed path within its own."""
    a = 0.5
    P = syn.thermal_splats(1, a, 72)
    _polarized_case(oracle, "thermal", P, "stored", res=12, N=128)
    cam = camera(15.0, 12)
    L = syn.length_unit()
    cache = make_cache(cam, 128, a, math.radians(60.0), "stored")
    ds = np.random.default_rng(31).normal(size=(12, 12, 1, 1, 4))
    g, cube = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds)
    gref, _ = oracle.polarized_gradient(a, math.radians(60.0), cam.alphas, cam.betas, 128, P, 0.0, 230e9, L,
                                        ds[:, :, 0, 0, :].reshape(cam.npix, 4, order="F"))
    assert np.abs(g - gref).max() <= 1e-9 * np.abs(gref).max()
    K.precision(cache, "mixed")
    g32, cube32 = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds)
    assert np.abs(cube32 - cube).max() <= 1e-4 * np.abs(cube[..., 0]).max()
    assert np.abs(g32 - g).max() <= 1e-3 * np.abs(g).max()


def _composite_parts(oracle, a, nt, npl, nk):
    T, P, Kp = syn.thermal_splats(1, a, nt), syn.powerlaw_splats(2, a, npl), syn.kappa_splats(3, a, nk)
    hyp = [oracle.kappa_hyp(Kp[13, i], math.exp(Kp[14, i])) for i in range(Kp.shape[1])]
    return [("thermal", T), ("powerlaw", P), ("kappa", Kp, hyp)]


def _composite_case(oracle, path, nt, npl, nk, res=16, N=256):
    """CompositeModel(PolarizedSplats, CompositeModel(PowerLawSplats, KappaSplats)) (transport.jl:298-316)
    against the oracle's composite at the FP64 image bars; the stored path also in mixed precision."""
    a, th = 0.9, math.radians(45.0)
    cam = camera(12.0, res)
    L = syn.length_unit()
    cache = make_cache(cam, N, a, th, path)
    ps = _composite_parts(oracle, a, nt, npl, nk)
    model = K.CompositeModel(K.Model(ps[0][1]), K.CompositeModel(K.Model(ps[1][1], kind="powerlaw"), K.Model(ps[2][1], kind="kappa")))
    assert model.ns == nt + npl + nk
    img = K.polarized_image(cache, model, 0.0, 230e9, L)
    ref = oracle.polarized_composite(a, th, cam.alphas, cam.betas, N, ps, 0.0, 230e9, L)[0].reshape(cam.size + (4,), order="F")
    I = np.abs(ref[..., 0])
    assert I.max() > 0
    err = np.abs(img - ref).max(axis=-1) / np.maximum(I, 1e-300)
    bright = I > 1e-6 * I.max()
    assert err[bright].max() < 1e-10, err[bright].max()
    assert np.abs(img - ref).max() <= 1e-10 * I.max()
    # every population reaches the image: the thermal part alone differs from the composite
    alone = K.polarized_image(cache, K.Model(ps[0][1]), 0.0, 230e9, L)
    assert np.abs(alone - img).max() > 1e-3 * I.max()
    if path == "stored":
        K.precision(cache, "mixed")
        got = K.polarized_image(cache, model, 0.0, 230e9, L)
        assert 0 < np.abs(got - img).max() <= 1e-4 * I.max()


@pytest.mark.parametrize("path", PATHS)
def test_polarized_composite_vs_oracle(oracle, path):
    _composite_case(oracle, path, 3, 2, 2)


def test_polarized_composite_tile_cull_vs_oracle(oracle):
    """72 splats of three populations: the transfer takes the per-tile cull with mixed population tags."""
    _composite_case(oracle, "stored", 40, 16, 16, res=12, N=128)


@pytest.mark.parametrize("path,nk1,nth", [("stored", 3, 3), ("fused", 3, 3), ("stored", 30, 40)])
def test_polarized_composite_knots_vs_oracle(oracle, path, nk1, nth):
    """CompositeModel(KnotSplats, PolarizedSplats, KnotSplats) with the two knot parts on different knot-time
    grids (each part's knots and times stay its own in the concatenated model) against the oracle's composite;
    with 72 splats the transfer would take the per-tile cull, which bounds pattern-rotating splats only and
    is left out for a composite holding knots."""
    from test_oracle_composite import knot_set
    a, th = 0.9, math.radians(45.0)
    cam = camera(12.0, 16)
    L = syn.length_unit()
    cache = make_cache(cam, 256, a, th, path)
    t1, t2 = np.array([-200.0, -100.0, 0.0, 100.0]), np.array([-150.0, -20.0, 60.0])
    P1, k1 = knot_set(a, nk1, t1, seed=4)
    P2, k2 = knot_set(a, 2, t2, seed=6)
    T = syn.thermal_splats(1, a, nth)
    ps = [("knot", P1, k1, t1), ("thermal", T), ("knot", P2, k2, t2)]
    model = K.CompositeModel(K.Model(P1, kind="knot", knots=k1, knot_times=t1),
                             K.CompositeModel(K.Model(T), K.Model(P2, kind="knot", knots=k2, knot_times=t2)))
    assert model.ns == nk1 + nth + 2
    for t_obs in (30.0, -170.0):
        img = K.polarized_image(cache, model, t_obs, 230e9, L)
        ref = oracle.polarized_composite(a, th, cam.alphas, cam.betas, 256, ps, t_obs, 230e9, L)[0].reshape(cam.size + (4,), order="F")
        I = np.abs(ref[..., 0])
        assert I.max() > 0
        assert np.abs(img - ref).max() <= 1e-10 * I.max()
        # the knot parts show: without both (the thermal part alone) and, with few splats, without the second
        for sub in [ps[1:2]] + ([ps[:2]] if nk1 < 10 else []):
            alone = oracle.polarized_composite(a, th, cam.alphas, cam.betas, 256, sub, t_obs, 230e9, L)[0]
            assert np.abs(alone.reshape(cam.size + (4,), order="F") - ref).max() > 1e-3 * I.max()
    if path == "stored":
        K.precision(cache, "mixed")
        got = K.polarized_image(cache, model, 30.0, 230e9, L)
        K.precision(cache, "fp64")
        img = K.polarized_image(cache, model, 30.0, 230e9, L)
        assert 0 < np.abs(got - img).max() <= 1e-4 * np.abs(img[..., 0]).max()


def test_composite_errors():
    th = K.Model(syn.thermal_splats(1, 0.9, 2))
    with pytest.raises(ValueError, match="parts"):
        K.CompositeModel(th, K.Model(syn.thin_splats(0, 0.9, 2), kind="thin"))
    with pytest.raises(ValueError):
        K.CompositeModel()
    cam = camera(10.0, 4)
    c = make_cache(cam, 16, 0.9, 0.5, "stored")
    thin = K.Model(syn.thin_splats(0, 0.9, 2), kind="thin")
    with pytest.raises(ValueError, match="polarized splats"):
        K.polarized_gradient(c, thin, [0.0], [230e9], 1e14, np.zeros((4, 4, 1, 1, 4)))
    # a composite's gradient is one vector of its parts' blocks
    with pytest.raises(K.api.DimensionMismatch):
        K.polarized_gradient(c, K.CompositeModel(th, th), [0.0], [230e9], 1e14, np.zeros((4, 4, 1, 1, 4)),
                             dparams=np.zeros((21, 4), order="F"))


@pytest.mark.parametrize("path", PATHS)
def test_cube_frames_match_images(oracle, path):
    """One launch over (frame, frequency) items equals per-frame images (polarized.jl:187-198)."""
    a, th = 0.9, math.radians(45.0)
    cam = camera(12.0, 12)
    L = syn.length_unit()
    P = syn.thermal_splats(2, a, 4, single_frame=False)
    cache = make_cache(cam, 128, a, th, path)
    m = K.Model(P)
    times, nus = [0.0, 20.0, 40.0], [86e9, 345e9]
    cube = K.polarized_cube(cache, m, times, nus, L)
    for it, t in enumerate(times):
        for inu, nu in enumerate(nus):
            img = K.polarized_image(cache, m, t, nu, L)
            assert np.array_equal(img, cube[:, :, it, inu, :])


def test_visibilities_vs_oracle(oracle):
    a, th = 0.5, math.radians(60.0)
    res = 16
    cam = camera(15.0, res)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, 4)
    cache = K.GeodesicCache(cam, 128, store_samples=False)
    K.regenerate(cache, a, th, marcher=K.Fused(64))
    img = K.polarized_image(cache, K.Model(P), 0.0, 230e9, L)
    u, v = syn.uv_disc(1, 64)
    psize = syn.pixel_size_rad(-15.0, 15.0, res)
    V = K.visibilities(cache, 1, psize, u, v)
    imgs = img.reshape(res * res, 4, order="F")[None]
    Vref = oracle.visibilities(imgs, res, res, psize, u, v)
    assert np.max(np.abs(V - Vref)) <= 1e-12 * np.max(np.abs(Vref))


def test_stored_and_fused_renders_agree():
    """The transfer from stored Recurrence(64) samples and the fused march + transfer see the same samples."""
    a, th = 0.5, math.radians(60.0)
    cam = camera(15.0, 16)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, 6)
    c1, c2 = make_cache(cam, 256, a, th, "stored"), make_cache(cam, 256, a, th, "fused")
    i1, i2 = K.polarized_image(c1, P, 0.0, 230e9, L), K.polarized_image(c2, P, 0.0, 230e9, L)
    assert np.abs(i1 - i2).max() <= 1e-13 * np.abs(i1[..., 0]).max()
    s1 = K.samples(c1)
    K.march(c1)
    s2 = K.samples(c1)
    for f in ("t", "r", "theta", "phi", "flags"):
        assert np.array_equal(s1[f], s2[f]), f


def _gradient_case(oracle, res=12, N=128, ns=3, dens=30.0, nmax=-1, slab=0.0, times=(0.0,), nus=(230e9,), seed=11):
    """polarized_gradient (GPU dual sweep over stored samples) against the oracle's dual sweep on the
    same seeded inputs: every entry within 1e-9 of the largest, and the forward image at the image bar.
    The device contracts the step adjoint with matrix-vector products and writes d ln G by hand where the
    oracle forms 4x4 dual matrices and 13-partial duals; per sample the two agree to about 7e-10 at
    degenerate spectra of K' (and 1e-12 elsewhere), hence 1e-9 rather than the image bar of 1e-10."""
    a, th = 0.5, math.radians(60.0)
    cam = camera(15.0, res)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, ns)
    P[12, :] += math.log(dens)
    cache = make_cache(cam, N, a, th, "stored")
    na, nb = cam.size
    ds = np.random.default_rng(seed).normal(size=(na, nb, len(times), len(nus), 4))
    g, cube = K.polarized_gradient(cache, P, list(times), list(nus), L, ds, nmax=nmax, slab=slab)
    gref = np.zeros_like(g)
    for it, t in enumerate(times):
        for inu, nu in enumerate(nus):
            dso = ds[:, :, it, inu, :].reshape(cam.npix, 4, order="F")
            gi, img = oracle.polarized_gradient(a, th, cam.alphas, cam.betas, N, P, t, nu, L, dso, nmax=nmax, slab=slab)
            gref += gi
            ref = img.reshape(cam.size + (4,), order="F")
            got = cube[:, :, it, inu, :]
            assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref[..., 0]).max()
    scale = np.abs(gref).max()
    assert scale > 0
    err = np.abs(g - gref).max() / scale
    assert err < 1e-9, err
    return g


def test_gradient_vs_oracle(oracle):
    _gradient_case(oracle)


def test_gradient_thin_parcels_vs_oracle(oracle):
    _gradient_case(oracle, dens=1.0, seed=12)


def test_gradient_winding_vs_oracle(oracle):
    _gradient_case(oracle, nmax=0, slab=0.5, seed=13)


def test_gradient_frames_and_frequencies_vs_oracle(oracle):
    _gradient_case(oracle, times=(0.0, 15.0), nus=(230e9, 345e9), seed=14)


def test_gradient_accumulates():
    """dparams accumulates (dual_gradient.jl:240: dparams .+= ...)."""
    a, th = 0.5, math.radians(60.0)
    cam = camera(15.0, 8)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, 2)
    cache = make_cache(cam, 64, a, th, "stored")
    ds = np.random.default_rng(3).normal(size=(8, 8, 1, 1, 4))
    g1, _ = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds)
    base = np.asfortranarray(np.full(P.shape, 2.5))
    g2, _ = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds, dparams=base)
    assert np.array_equal(g2, g1 + 2.5)


def test_errors_mirror_reference():
    cam = camera(10.0, 4)
    c = K.GeodesicCache(cam, 8, store_samples=False)
    with pytest.raises(ValueError, match="without sample storage"):
        K.regenerate(c, 0.9, 0.5, marcher=K.Direct())
    with pytest.raises(ValueError, match="regenerate"):
        K.polarized_image(c, syn.thermal_splats(1, 0.9, 1), 0.0, 230e9, 1e14)
    with pytest.raises(K.api.DimensionMismatch):
        K.regenerate(c, 0.9, 0.5, camera=camera(10.0, 5), marcher=K.Fused(64))
    with pytest.raises(ValueError):
        K.GeodesicCache(cam, 0)
    K.regenerate(c, 0.9, 0.5, marcher=K.Fused(64))
    with pytest.raises(ValueError, match="no stored samples"):
        K.polarized_gradient(c, syn.thermal_splats(1, 0.9, 1), [0.0], [230e9], 1e14, np.zeros((4, 4, 1, 1, 4)))


# ---- mixed precision (FP64 geodesics, FP32 transfer and adjoint) ------------------------------------------------
# Bars from the target of the mixed path: Stokes error <= 1e-4 of the peak intensity and gradient relative
# error <= 1e-3 of the largest entry, both against the plain FP64 path on identical inputs (itself at parity
# with the oracle above).
MIXED_CASES = [("thermal", syn.thermal_splats(1, 0.5, 6), 0.5, 60.0, 15.0, {}),
               ("powerlaw", syn.powerlaw_splats(2, 0.9, 4), 0.9, 45.0, 12.0, {}),
               ("kappa", syn.kappa_splats(2, 0.9, 4), 0.9, 45.0, 12.0, {})]


@pytest.mark.parametrize("kind,P,a,thdeg,fov,kw", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mixed_image_error(kind, P, a, thdeg, fov, kw):
    th = math.radians(thdeg)
    cam = camera(fov, 24)
    L = syn.length_unit()
    cache = make_cache(cam, 512, a, th, "stored")
    m = K.Model(P, kind=kind, **kw)
    ref = K.polarized_cube(cache, m, [0.0, 20.0], [230e9, 345e9], L)
    K.precision(cache, "mixed")
    got = K.polarized_cube(cache, m, [0.0, 20.0], [230e9, 345e9], L)
    for it in range(2):
        for inu in range(2):
            peak = np.abs(ref[:, :, it, inu, 0]).max()
            assert peak > 0
            err = np.abs(got[:, :, it, inu] - ref[:, :, it, inu]).max() / peak
            assert err <= 1e-4, (kind, it, inu, err)
            assert err > 0          # the FP32 path really ran


def test_mixed_dense_parcels_error():
    """Optically thick, Faraday-rotating parcels: the FP64 fallback of the step keeps the FP32 path at its bar."""
    a, th = 0.5, math.radians(60.0)
    cam = camera(15.0, 24)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, 6)
    P[12, :] += math.log(300.0)
    cache = make_cache(cam, 512, a, th, "stored")
    ref = K.polarized_image(cache, P, 0.0, 230e9, L)
    K.precision(cache, "mixed")
    got = K.polarized_image(cache, P, 0.0, 230e9, L)
    err = np.abs(got - ref).max() / np.abs(ref[..., 0]).max()
    assert 0 < err <= 1e-4, err


@pytest.mark.parametrize("dens", [1.0, 30.0])
def test_mixed_gradient_error(dens):
    a, th = 0.5, math.radians(60.0)
    cam = camera(15.0, 16)
    L = syn.length_unit()
    P = syn.thermal_splats(1, a, 4)
    P[12, :] += math.log(dens)
    cache = make_cache(cam, 256, a, th, "stored")
    ds = np.random.default_rng(21).normal(size=(16, 16, 1, 1, 4))
    g64, img64 = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds)
    K.precision(cache, "mixed")
    g32, img32 = K.polarized_gradient(cache, P, [0.0], [230e9], L, ds)
    gerr = np.abs(g32 - g64).max() / np.abs(g64).max()
    ierr = np.abs(img32 - img64).max() / np.abs(img64[..., 0]).max()
    assert 0 < gerr <= 1e-3, gerr
    assert ierr <= 1e-4, ierr


@pytest.mark.parametrize("a,thdeg,fov", [(0.94, 17.0, 12.0), (0.5, 60.0, 15.0), (0.9, 163.0, 10.0)])
def test_pixel_constants_and_residuals_vs_oracle(oracle, a, thdeg, fov):
    """host(cache.consts) and the per-ray anchor residuals (cache.jl:98-99, quadrature.jl:283-291) against the
    oracle's K1 and Recurrence(64) march: the root count and case bit-exact, the per-pixel constants to 1e-10
    (scaled by max(|x|, 1)); the residuals are the largest differences between the recurrence and the closed
    forms at the anchors, rounding-level in both, and agree to 1e-9."""
    th = math.radians(thdeg)
    cam = camera(fov, 16)
    N = 256
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, a, th, marcher=K.Recurrence(64))
    c = K.consts(cache)
    idx = cache.perm_host - 1
    assert np.array_equal(c["alpha"], cam.alphas[idx]) and np.array_equal(c["beta"], cam.betas[idx])
    for j in range(cache.npix):
        ref = oracle.pixel_constants(a, th, cam.alphas[idx[j]], cam.betas[idx[j]])
        assert int(c["nreal"][j]) == ref["nreal"] and int(c["case"][j]) == ref["case"], j
        for f in ("tau_total", "fo", "k", "lam", "eta", "up", "um"):
            if np.isfinite(ref[f]) or np.isfinite(c[f][j]):
                assert scaled_err(c[f][j], ref[f]) < 1e-10, (j, f, c[f][j], ref[f])
    rt, rp = K.residuals(cache)
    smp = oracle.samples(a, th, cam.alphas[idx], cam.betas[idx], N, marcher=1)
    assert np.all(np.isfinite(rt)) and np.all(np.isfinite(rp))
    assert np.abs(rt - smp["resid_t"]).max() < 1e-9 and np.abs(rp - smp["resid_phi"]).max() < 1e-9
    assert np.abs(rt).max() < 2e-8 and np.abs(rp).max() < 2e-9


def test_reflection_parity_on_the_device(oracle):
    """The equatorial and azimuthal reflections of the reference's parity gate (test_polarized_splats.jl:197-241)
    on the device: theta_o = 163 deg is the beta-mirror of the reflected source at 17 deg, spin -a the
    alpha-mirror at +a, with U and V reversed (1e-10); and the 163 deg image against the oracle's (1e-10 of
    each pixel's I)."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle_endtoend import reflection_params, rel, grav
    res, N, fov = 24, 60, 20.0
    d = fov / res
    cam = K.Camera((-fov / 2 + d / 2, fov / 2 - d / 2), (-fov / 2 + d / 2, fov / 2 - d / 2), res)
    L = grav(6.5e9)

    def image(a, th, P):
        cache = K.GeodesicCache(cam, N, store_samples=False)
        K.regenerate(cache, a, th, marcher=K.Fused(64))
        return K.polarized_image(cache, P, 0.0, 230e9, L)

    p = reflection_params()
    A = image(0.94, math.radians(163.0), p)
    q = p.copy(); q[2] = -q[2]; q[16] += math.pi; q[19] = -q[19]
    D = image(0.94, math.radians(17.0), q)
    sgn = np.array([1, 1, -1, -1])
    for k in range(4):
        assert rel(D[:, ::-1, k] * sgn[k], A[:, :, k]) < 1e-10, k
    B = image(0.94, math.radians(17.0), p)
    r = p.copy(); r[1] = -r[1]; r[15] = math.pi - r[15]; r[16] = math.pi - r[16]; r[18] = -r[18]
    E = image(-0.94, math.radians(17.0), r)
    for k in range(4):
        assert rel(E[::-1, :, k] * sgn[k], B[:, :, k]) < 1e-10, k
    ref, _, _ = oracle.polarized_image(0.94, math.radians(163.0), cam.alphas, cam.betas, N, p, 0.0, 230e9, L)
    ref = ref.reshape(res, res, 4, order="F")
    Iref = np.maximum(np.abs(ref[..., 0]), 1e-300)
    assert (np.abs(A - ref).max(axis=2) / Iref)[ref[..., 0] > 1e-6 * ref[..., 0].max()].max() < 1e-10
