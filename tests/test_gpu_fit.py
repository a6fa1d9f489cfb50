"""GPU parity of the likelihood layer: visibilities, the χ² of a frame and its adjoint seed computed on
the device, the two passes of the dual sweep, and the movie χ² gradient.

The oracle's visibilities and seeds are pinned separately (tests/test_oracle_visibilities.py: ehtim's
observation of a known source, the adjoint identity, central differences of χ²). Here the device
results are compared with them on the same seeded inputs: visibilities and χ² to 1e-12 relative
(FP64 GEMMs over separable phase tables against a direct sum), seeds through the gradient they
produce against the host-seeded dual sweep at the gradient bar (1e-9 of the largest entry), and the
movie χ² against the host sum of the rendered cube.
"""
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

A, THO = 0.5, math.radians(60.0)
NU = 230e9


def stored_cache(res, N, fov=15.0, nb=None):
    cam = K.Camera((-fov, fov), (-fov, fov), res) if nb is None else K.Camera((-fov, fov), (-fov * nb / res, fov * nb / res), (res, nb))
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    return cam, cache


def screen_flat(x):
    """(n_alpha, n_beta, 4) -> (npix, 4) with pixel = i + n_alpha j."""
    return np.transpose(x, (1, 0, 2)).reshape(-1, 4)


@pytest.mark.parametrize("res,nuv,nframes", [(16, 64, 1), (24, 200, 3)])
def test_visibilities_gemm_vs_oracle(oracle, res, nuv, nframes):
    cam, cache = stored_cache(res, 128)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 4)
    times = [0.0, 30.0, 60.0][:nframes]
    cube = K.polarized_cube(cache, P, times, [NU], L)
    u, v = syn.uv_disc(1, nuv)
    tap = np.random.default_rng(5).uniform(0.5, 1.0, nuv)
    psize = syn.pixel_size_rad(-15.0, 15.0, res)
    V = K.visibilities(cache, nframes, psize, u, v, taper=tap)
    imgs = np.stack([screen_flat(cube[:, :, f, 0, :]) for f in range(nframes)])
    Vref = oracle.visibilities(imgs, res, res, psize, u, v, taper=tap)
    assert np.abs(V - Vref).max() <= 1e-12 * np.abs(Vref).max()


def _vis_data(oracle, img_screen, na, nb, psize, nuv, seed):
    rng = np.random.default_rng(seed)
    u, v = syn.uv_disc(1, nuv)
    V0 = oracle.visibilities(img_screen[None], na, nb, psize, u, v)[0]
    sig = 0.02 * np.abs(V0[:, 0]).max() * rng.uniform(0.5, 1.5, size=(nuv, 4))
    d = V0 * (1 + 0.1 * rng.normal(size=V0.shape)) + sig * rng.normal(size=V0.shape)
    g = np.exp(rng.normal(scale=0.05, size=nuv) + 1j * rng.normal(scale=0.2, size=nuv))
    tap = rng.uniform(0.6, 1.0, nuv)
    return u, v, d, sig, g, tap


@pytest.mark.parametrize("na,nb", [(16, 16), (20, 12)])
def test_vis_chi2_seed_vs_oracle(oracle, na, nb):
    """frame_chi2_seed!: χ², the model visibilities, and the gradient the device seed produces."""
    cam, cache = stored_cache(na, 128, nb=nb)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(30.0)
    m = K.Model(P)
    img_sorted = K.gradient_forward(cache, m, [0.0], [NU], L)[0, 0]
    img = K.unsort(cache, img_sorted)                        # screen order
    psize = syn.pixel_size_rad(-15.0, 15.0, na)
    u, v, d, sig, g, tap = _vis_data(oracle, img, na, nb, psize, 96, 7)
    chi, V = K.vis_chi2_seed(cache, 0, psize, u, v, d, sig, taper=tap, gains=g, return_model=True)
    Vref = oracle.visibilities(img[None], na, nb, psize, u, v, taper=tap)[0]
    assert np.abs(V - Vref).max() <= 1e-12 * np.abs(Vref).max()
    r = (g[:, None] * Vref - d) / sig
    chi_ref = float(np.sum(np.abs(r) ** 2))
    assert abs(chi - chi_ref) <= 1e-11 * chi_ref
    dp_dev = np.zeros(P.shape, order="F")
    K.gradient_reverse(cache, m, dp_dev)                      # the device seed
    w = 2 * np.conj(g)[:, None] * r / sig
    seed = oracle.vis_seed(w[None], na, nb, psize, u, v, taper=tap)[0]          # screen order
    ds = np.transpose(seed.reshape(nb, na, 4), (1, 0, 2))[:, :, None, None, :]
    dp_host, _ = K.polarized_gradient(cache, m, [0.0], [NU], L, ds)
    scale = np.abs(dp_host).max()
    assert scale > 0
    assert np.abs(dp_dev - dp_host).max() <= 1e-9 * scale


def test_image_chi2_seed_and_movie_gradient(oracle):
    """_frame_seed! on the device (per-pixel noise, a mask) and chi2_gradient over a two-frame movie
    against the host χ² of the rendered cube and the host-seeded dual sweep."""
    na = 12
    cam, cache = stored_cache(na, 128)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(30.0)
    m = K.Model(P)
    times = np.array([0.0, 25.0])
    truth = K.polarized_cube(cache, P * 1.0 + 0.02, times, [NU], L)             # perturbed model as data
    rng = np.random.default_rng(9)
    sigma = 0.05 * np.abs(truth[..., 0]).max() * rng.uniform(0.5, 1.5, size=truth.shape)
    data = truth + sigma * rng.normal(size=truth.shape)
    mask = rng.uniform(size=truth.shape[:4]) > 0.2
    movie = K.StokesMovie(data, times, [NU], sigma, mask)
    dp = np.zeros(P.shape, order="F")
    chi = K.chi2_gradient(dp, m, movie, cache, L, batch_frames=2)
    chi_ref = K.chi2(m, movie, cache, L)
    assert abs(chi - chi_ref) <= 1e-12 * chi_ref
    cube = K.polarized_cube(cache, m, times, [NU], L)
    r = (cube - data) / sigma
    seed = np.where(mask[..., None], 2 * r / sigma, 0.0)
    dp_ref, _ = K.polarized_gradient(cache, m, list(times), [NU], L, seed)
    assert np.abs(dp - dp_ref).max() <= 1e-9 * np.abs(dp_ref).max()
    # one frame at a time gives the same sums
    dp1 = np.zeros(P.shape, order="F")
    chi1 = K.chi2_gradient(dp1, m, movie, cache, L, batch_frames=1)
    assert abs(chi1 - chi) <= 1e-12 * chi
    assert np.abs(dp1 - dp).max() <= 1e-12 * np.abs(dp).max()


def test_binned_chi2_seed_and_movie_gradient():
    """_frame_seed! through a Binning on the device (2 x 2 points per pixel) against the host restatement
    (tests/test_binning.py, pinned by central differences): the movie χ² against the host χ² of the binned
    cube, the gradient against the dual sweep seeded on the host; image_loss_gradient through the binning."""
    from test_binning import frame_seed_binned
    cam, b = K.binned_grid((-15.0, 15.0), (-15.0, 15.0), (8, 8), subsamples=2)
    cache = K.GeodesicCache(cam, 128, store_samples=True)
    K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(30.0)
    m = K.Model(P)
    times = np.array([0.0, 25.0])
    truth = K.bin(b, K.polarized_cube(cache, P * 1.0 + 0.02, times, [NU], L))       # (8, 8, 2, 1, 4)
    rng = np.random.default_rng(12)
    sigma = 0.05 * np.abs(truth[..., 0]).max() * rng.uniform(0.5, 1.5, size=truth.shape)
    data = truth + sigma * rng.normal(size=truth.shape)
    mask = rng.uniform(size=truth.shape[:4]) > 0.2
    movie = K.StokesMovie(data, times, [NU], sigma, mask)
    dp = np.zeros(P.shape, order="F")
    chi = K.chi2_gradient(dp, m, movie, cache, L, batch_frames=2, binning=b)
    chi_ref = K.chi2(m, movie, cache, L, binning=b)
    assert abs(chi - chi_ref) <= 1e-12 * chi_ref
    cube = K.polarized_cube(cache, m, times, [NU], L)                                 # (256, 1, 2, 1, 4)
    seeds = np.zeros(cube.shape)
    chi_host = 0.0
    for it in range(2):
        c_, s_ = frame_seed_binned(cube[:, 0, it, 0], screen_flat(data[:, :, it, 0]), screen_flat(sigma[:, :, it, 0]),
                                   mask[:, :, it, 0].T.ravel(), b)
        seeds[:, 0, it, 0] = s_
        chi_host += c_
    assert abs(chi - chi_host) <= 1e-12 * chi_host
    dp_ref, _ = K.polarized_gradient(cache, m, list(times), [NU], L, seeds)
    assert np.abs(dp - dp_ref).max() <= 1e-9 * np.abs(dp_ref).max()
    # a linear loss of the binned image: its seed through the binning is w[pixel] / count per point
    w = rng.normal(size=(8, 8, 4))
    dl = np.zeros(P.shape, order="F")
    val = K.image_loss_gradient(dl, lambda img: (float(np.sum(w * img)), w), cache, m, 0.0, NU, L, binning=b)
    assert abs(val - np.sum(w * K.bin(b, cube[:, :, 0, 0]))) <= 1e-12 * abs(val)
    pw = screen_flat(w)[b.pixel] / b.count[b.pixel][:, None]
    dl_ref, _ = K.polarized_gradient(cache, m, [0.0], [NU], L, pw[:, None, None, None, :])
    assert np.abs(dl - dl_ref).max() <= 1e-9 * np.abs(dl_ref).max()


def test_binning_of_one_point_per_pixel_is_the_unbinned_seed():
    na = 10
    cam, cache = stored_cache(na, 96)
    L = syn.length_unit()
    m = K.Model(syn.thermal_splats(1, A, 2))
    img = K.polarized_cube(cache, m, [0.0], [NU], L)[:, :, 0, 0]
    rng = np.random.default_rng(4)
    data, sigma = img * rng.uniform(0.9, 1.1, size=img.shape), 0.1 * np.abs(img[..., 0]).max() * np.ones(4)
    K.gradient_forward(cache, m, [0.0], [NU], L, copy=False)
    b = K.Binning(np.arange(na * na), (na, na))
    chi_b = K.image_chi2_seed(cache, 0, data, sigma, binning=b)
    g_b = np.zeros((21, 2), order="F")
    K.gradient_reverse(cache, m, g_b, None)
    chi = K.image_chi2_seed(cache, 0, data, sigma)
    g = np.zeros((21, 2), order="F")
    K.gradient_reverse(cache, m, g, None)
    g2 = np.zeros((21, 2), order="F")
    K.gradient_reverse(cache, m, g2, None)                         # reverse! again on the same forward pass
    assert np.all(np.isfinite(g)) and np.array_equal(g2, g)
    assert abs(chi_b - chi) <= 1e-13 * chi                       # r / sigma against r * (1 / sigma): rounding only
    assert np.abs(g_b - g).max() <= 1e-12 * np.abs(g).max()
    with pytest.raises(K.api.DimensionMismatch, match="points"):
        K.image_chi2_seed(cache, 0, data[:5, :5], sigma, binning=K.Binning(np.arange(25), (5, 5)))


def test_forward_reverse_equal_polarized_gradient():
    na = 10
    cam, cache = stored_cache(na, 96)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 2)
    m = K.Model(P)
    ds = np.random.default_rng(2).normal(size=(na, na, 1, 1, 4))
    g_ref, cube = K.polarized_gradient(cache, m, [0.0], [NU], L, ds)
    dp = np.zeros(P.shape, order="F")
    fwd, rev = K.sweep_passes(dp, cache, m, L)
    img = fwd(0.0, NU)
    assert np.array_equal(K.unsort(cache, img), screen_flat(cube[:, :, 0, 0, :]))
    ds_sorted = screen_flat(ds[:, :, 0, 0, :])[cache.perm_host - 1]
    rev(ds_sorted, 0.0, NU)
    assert np.array_equal(dp, g_ref)


def test_seed_errors():
    na = 6
    cam, cache = stored_cache(na, 64)
    L = syn.length_unit()
    m = K.Model(syn.thermal_splats(1, A, 1))
    dp = np.zeros(m.params.shape, order="F")
    with pytest.raises(ValueError, match="forward"):
        K.gradient_reverse(cache, m, dp)
    K.gradient_forward(cache, m, [0.0, 10.0], [NU], L, copy=False)
    with pytest.raises(ValueError, match="no device seed"):
        K.gradient_reverse(cache, m, dp)
    with pytest.raises(ValueError, match="outside"):
        K.image_chi2_seed(cache, 2, np.zeros((na, na, 4)), np.ones(4))
    with pytest.raises(ValueError, match="sigma"):
        K.image_chi2_seed(cache, 0, np.zeros((na, na, 4)), np.array([1.0, 0.0, 1.0, 1.0]))
    other = K.Model(syn.thermal_splats(1, A, 1))
    with pytest.raises(ValueError, match="model of its forward"):
        K.gradient_reverse(cache, other, dp, np.zeros((1, 2, cache.npix, 4)))
    K.gradient_reverse(cache, m, dp, np.zeros((1, 2, cache.npix, 4)))     # host seeds replace the device seeds
    with pytest.raises(ValueError, match="no device seed"):
        K.gradient_reverse(cache, m, dp)
    K.polarized_image(cache, m, 0.0, NU, L)                               # a render ends the forward pass
    with pytest.raises(K.api.StateError, match="forward"):
        K.gradient_reverse(cache, m, dp, np.zeros((1, 2, cache.npix, 4)))


def test_reverse_rejects_a_model_recreated_at_the_same_address():
    """A model freed after the forward pass and a new one created (often at the freed address) is not the
    forward pass's model: the library tells models apart by a creation id, not by their address."""
    na = 6
    cam, cache = stored_cache(na, 64)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 1)
    m = K.Model(P)
    K.gradient_forward(cache, m, [0.0], [NU], L, copy=False)
    old = m._h
    m.__del__()
    same_address = False
    for _ in range(64):
        m2 = K.Model(P)
        if m2._h == old:
            same_address = True
            break
    print(f"recreated at the freed address: {same_address}", flush=True)
    dp = np.zeros(P.shape, order="F")
    with pytest.raises(ValueError, match="model of its forward"):
        K.gradient_reverse(cache, m2, dp, np.zeros((1, 1, cache.npix, 4)))


@pytest.mark.parametrize("on_host", [False, True])
def test_closure_seed_vs_oracle(oracle, on_host):
    """frame_chi2_closures: closure χ² of a frame's visibilities and its seed, added to a visibility χ² seed of the
    same frame, on the device (ks_closure_chi2_seed: closure phases and log closure amplitudes, their χ² and
    dχ²/dV, the adjoint transform) or with the closures on the host and the seed through vis_seed; the χ² equals
    the host formula on the oracle's visibilities and the gradient it drives equals the host-seeded dual sweep with
    the oracle's adjoint transform."""
    import itertools
    from kerrsplat_amd import fit
    na = 14
    cam, cache = stored_cache(na, 128)
    L = syn.length_unit()
    P = syn.thermal_splats(1, A, 3)
    P[12, :] += math.log(30.0)
    m = K.Model(P)
    img = K.unsort(cache, K.gradient_forward(cache, m, [0.0], [NU], L)[0, 0])
    psize = syn.pixel_size_rad(-15.0, 15.0, na)
    rng = np.random.default_rng(13)
    nst = 6
    pairs = list(itertools.combinations(range(1, nst + 1), 2))
    index = {p: n + 1 for n, p in enumerate(pairs)}
    leg = lambda a, b: index[(a, b)] if a < b else -index[(b, a)]
    pos = rng.normal(scale=3e9, size=(nst, 2))
    u = np.array([pos[i - 1, 0] - pos[j - 1, 0] for i, j in pairs]); v = np.array([pos[i - 1, 1] - pos[j - 1, 1] for i, j in pairs])
    tri = [(leg(i, j), leg(j, k), leg(k, i)) for i, j, k in itertools.combinations(range(1, nst + 1), 3)]
    quad = [(leg(i, j), leg(k, l), leg(i, k), leg(j, l)) for i, j, k, l in itertools.combinations(range(1, nst + 1), 4)]
    V0 = oracle.visibilities(img[None], na, na, psize, u, v)[0]
    data = fit.ClosureData(u, v, tri, fit.closure_phases(V0, tri) + rng.normal(scale=0.2, size=len(tri)), np.full(len(tri), 0.1),
                           quad, fit.log_closure_amplitudes(V0, quad) + rng.normal(scale=0.1, size=len(quad)), np.full(len(quad), 0.1))
    sig = 0.05 * np.abs(V0[:, 0]).max() * np.ones((len(u), 4))
    vdata = V0 * 1.1
    chi_v = K.vis_chi2_seed(cache, 0, psize, u, v, vdata, sig)
    dalpha = 30.0 / (na - 1)
    chi_c = fit.frame_chi2_closures(cache, 0, data, dalpha, L, syn.DISTANCE_CM, add=True, on_host=on_host)
    chi_c_ref, w = fit.closure_chi2(V0, data, grad=True)
    assert abs(chi_c - chi_c_ref) <= 1e-9 * chi_c_ref
    dp_dev = np.zeros(P.shape, order="F")
    K.gradient_reverse(cache, m, dp_dev)
    wv = 2 * (V0 - vdata) / sig ** 2
    seed = oracle.vis_seed((w + wv)[None], na, na, psize, u, v)[0]
    ds = np.transpose(seed.reshape(na, na, 4), (1, 0, 2))[:, :, None, None, :]
    dp_host, _ = K.polarized_gradient(cache, m, [0.0], [NU], L, ds)
    assert chi_v > 0 and np.abs(dp_host).max() > 0
    assert np.abs(dp_dev - dp_host).max() <= 1e-9 * np.abs(dp_host).max()
