"""The forward + adjoint chi2 gradient workload (a = 0.7, theta_o = 30 deg, 64 thermal splats = 1344 gradient rows,
4 frames over 30 M, 230 GHz; data rendered from the parameters perturbed by N(0, 5%) plus noise at 1% of the
peak; losses: the image chi2, the visibility chi2 on 2048 uv points and closure phases over 10 stations).

* On a reduced screen (20 x 20, N = 256): the device gradient of the whole loss (image chi2 seeds, visibility chi2
  and closure seeds added on the device) against the oracle's dual sweep driven by the same seeds computed on
  the host from the oracle's images, at the FP64 gradient bar 1e-9 of the largest entry; and against central
  finite differences of the loss evaluated from device renders (relative step 1e-6) for a set of rows.
* At the full size (128 x 128, N = 1024, 4 frames): the mixed path against the FP64 path (images 1e-4 of the
  peak, gradient 1e-3 of the largest entry) and the gradient linear in the seed.
The visibility and closure uv sets are separate (2048 points uniform in a disc; the 45 baselines of 10 stations).
"""
import itertools
import math

import numpy as np
import pytest

import kerrsplat_amd as K
from kerrsplat_amd import fit
from kerrsplat_amd import synthetic as syn

pytestmark = pytest.mark.gpu

C = syn.CONFIGS[3]
A, THO, NU, TIMES = C["a"], math.radians(C["theta_o_deg"]), C["nus"][0], C["times"]


def closure_setup(nst=10):
    pos = syn.stations(3, nst)
    pairs = list(itertools.combinations(range(1, nst + 1), 2))
    index = {p: n + 1 for n, p in enumerate(pairs)}
    leg = lambda a, b: index[(a, b)] if a < b else -index[(b, a)]
    u = np.array([pos[i - 1, 0] - pos[j - 1, 0] for i, j in pairs])
    v = np.array([pos[i - 1, 1] - pos[j - 1, 1] for i, j in pairs])
    tri = [(leg(i, j), leg(j, k), leg(k, i)) for i, j, k in itertools.combinations(range(1, nst + 1), 3)]
    return u, v, tri


class Problem:
    """The cfg3 likelihood on a res x res screen with N samples: data from perturbed parameters."""

    def __init__(self, oracle, res, N):
        self.oracle = oracle
        _, P = syn.config_model(3)
        self.P = P
        self.res, self.N = res, N
        self.cam = K.Camera((-C["fov"], C["fov"]), (-C["fov"], C["fov"]), res)
        self.L = syn.length_unit()
        self.psize = syn.pixel_size_rad(-C["fov"], C["fov"], res)
        rng = np.random.default_rng(33)
        Pd = P * (1.0 + 0.05 * rng.normal(size=P.shape))
        cache = K.GeodesicCache(self.cam, N, store_samples=True)
        K.regenerate(cache, A, THO, marcher=K.Recurrence(64))
        self.cache = cache
        truth = K.polarized_cube(cache, K.Model(Pd), TIMES, [NU], self.L)[:, :, :, 0, :]      # (res, res, nt, 4)
        peak = np.abs(truth[..., 0]).max()
        self.sig_img = np.full(4, 0.01 * peak)
        self.data_img = truth + self.sig_img * rng.normal(size=truth.shape)
        self.u, self.v = syn.uv_disc(3, C["nuv"])
        self.cu, self.cv, self.tri = closure_setup(C["stations"])
        self.Vd, self.sig_v, self.cdata = [], [], []
        for f in range(len(TIMES)):
            img = truth[:, :, f, :].reshape(-1, 4, order="F")
            V0 = oracle.visibilities(img[None], res, res, self.psize, self.u, self.v)[0]
            sv = 0.01 * np.abs(V0[:, 0]).max() * np.ones((len(self.u), 4))
            self.Vd.append(V0 + sv * (rng.normal(size=V0.shape) + 1j * rng.normal(size=V0.shape)))
            self.sig_v.append(sv)
            Vc = oracle.visibilities(img[None], res, res, self.psize, self.cu, self.cv)[0]
            cp = fit.closure_phases(Vc, self.tri) + rng.normal(scale=0.05, size=len(self.tri))
            self.cdata.append(fit.ClosureData(self.cu, self.cv, self.tri, cp, np.full(len(self.tri), 0.05)))

    def loss_from_cube(self, cube):
        """The whole loss of rendered frames (res, res, nt, 4), evaluated on the host."""
        total = 0.0
        for f in range(len(TIMES)):
            r = (cube[:, :, f, :] - self.data_img[:, :, f, :]) / self.sig_img
            total += float(np.sum(r ** 2))
            img = cube[:, :, f, :].reshape(-1, 4, order="F")[None]
            ovis = self.oracle.visibilities
            V = ovis(img, self.res, self.res, self.psize, self.u, self.v)[0]
            total += float(np.sum(np.abs((V - self.Vd[f]) / self.sig_v[f]) ** 2))
            Vc = ovis(img, self.res, self.res, self.psize, self.cu, self.cv)[0]
            total += fit.closure_chi2(Vc, self.cdata[f])
        return total

    def device_gradient(self, model):
        """chi2 and its gradient from the device: forward pass, per frame the image chi2 seed, the visibility chi2
        seed (dchi2/dV through vis_seed, added) and the closure seed (added), one reverse pass."""
        cache = self.cache
        dp = np.zeros(model.grad_shape, order="F")
        K.gradient_forward(cache, model, TIMES, [NU], self.L, copy=False)
        total = 0.0
        V = K.visibilities(cache, len(TIMES), self.psize, self.u, self.v)
        dalpha = 2 * C["fov"] / (self.res - 1)
        for f in range(len(TIMES)):
            total += K.image_chi2_seed(cache, f, self.data_img[:, :, f, :], self.sig_img)
            rv = V[f] - self.Vd[f]
            total += float(np.sum(np.abs(rv / self.sig_v[f]) ** 2))
            K.vis_seed(cache, f, self.psize, self.u, self.v, 2 * rv / self.sig_v[f] ** 2, add=True)
            total += fit.frame_chi2_closures(cache, f, self.cdata[f], dalpha, self.L, syn.DISTANCE_CM, add=True)
        K.gradient_reverse(cache, model, dp)
        return total, dp


def test_cfg3_gradient_vs_oracle_and_central_differences(oracle):
    res, N = 20, 256
    pb = Problem(oracle, res, N)
    m = K.Model(pb.P)
    print(f"cfg3 reduced: {res}x{res} x {N}, 64 thermal splats, 4 frames", flush=True)
    chi, g = pb.device_gradient(m)
    # the oracle: its images, the same seeds on the host, its dual sweep per frame
    gref = np.zeros_like(g)
    total = 0.0
    cam = pb.cam
    for f, t in enumerate(TIMES):
        img, _, _ = oracle.polarized_image(A, THO, cam.alphas, cam.betas, N, pb.P, t, NU, pb.L)
        scr = img.reshape(res, res, 4, order="F")
        r = (scr - pb.data_img[:, :, f, :]) / pb.sig_img
        total += float(np.sum(r ** 2))
        seed = (2 * r / pb.sig_img).reshape(-1, 4, order="F")
        V = oracle.visibilities(img[None], res, res, pb.psize, pb.u, pb.v)[0]
        rv = V - pb.Vd[f]
        total += float(np.sum(np.abs(rv / pb.sig_v[f]) ** 2))
        seed += oracle.vis_seed((2 * rv / pb.sig_v[f] ** 2)[None], res, res, pb.psize, pb.u, pb.v)[0]
        Vc = oracle.visibilities(img[None], res, res, pb.psize, pb.cu, pb.cv)[0]
        cchi, wc = fit.closure_chi2(Vc, pb.cdata[f], grad=True)
        total += cchi
        seed += oracle.vis_seed(wc[None], res, res, pb.psize, pb.cu, pb.cv)[0]
        gf, _ = oracle.polarized_gradient(A, THO, cam.alphas, cam.betas, N, pb.P, t, NU, pb.L, seed)
        gref += gf
    assert abs(chi - total) <= 1e-9 * total
    gmax = np.abs(gref).max()
    assert gmax > 0
    assert np.abs(g - gref).max() <= 1e-9 * gmax, np.abs(g - gref).max() / gmax
    # central differences of the loss from device renders, relative step 1e-6, on the largest rows
    rows = np.argsort(np.abs(g).ravel(order="F"))[::-1][:6]
    for k in rows:
        r_, i_ = k % g.shape[0], k // g.shape[0]
        h = 1e-6 * max(abs(pb.P[r_, i_]), 0.05)
        Qp, Qm = pb.P.copy(), pb.P.copy()
        Qp[r_, i_] += h
        Qm[r_, i_] -= h
        lp = pb.loss_from_cube(K.polarized_cube(pb.cache, K.Model(Qp), TIMES, [NU], pb.L)[:, :, :, 0, :])
        lm = pb.loss_from_cube(K.polarized_cube(pb.cache, K.Model(Qm), TIMES, [NU], pb.L)[:, :, :, 0, :])
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[r_, i_]) <= 1e-6 * gmax, (r_, i_, fd, g[r_, i_])


def test_cfg3_full_size_mixed_and_linearity(oracle):
    res, N = C["res"], C["N"]
    pb = Problem(oracle, res, N)
    m = K.Model(pb.P)
    print(f"cfg3 full size: {res}x{res} x {N}, 4 frames", flush=True)
    rng = np.random.default_rng(5)
    ds = rng.normal(size=(res, res, len(TIMES), 1, 4))
    g64, cube64 = K.polarized_gradient(pb.cache, m, TIMES, [NU], pb.L, ds)
    g64b, _ = K.polarized_gradient(pb.cache, m, TIMES, [NU], pb.L, 2.0 * ds)
    assert np.abs(g64b - 2.0 * g64).max() <= 1e-12 * np.abs(g64).max()
    K.precision(pb.cache, "mixed")
    try:
        g32, cube32 = K.polarized_gradient(pb.cache, m, TIMES, [NU], pb.L, ds)
        chi32, gc32 = pb.device_gradient(m)
    finally:
        K.precision(pb.cache, "fp64")
    chi64, gc64 = pb.device_gradient(m)
    peak = np.abs(cube64[..., 0]).max()
    assert np.abs(cube32 - cube64).max() <= 1e-4 * peak
    assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()
    assert abs(chi32 - chi64) <= 1e-3 * chi64
    assert np.abs(gc32 - gc64).max() <= 1e-3 * np.abs(gc64).max()
