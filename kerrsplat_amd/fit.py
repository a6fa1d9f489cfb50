"""The likelihood side of the hot path, mirroring src/Fit/Fit.jl and src/Fit/device_visibilities.jl:
a Stokes movie, its χ² and the χ² gradient through the device dual sweep, and the visibility χ² of a
frame with its adjoint seed computed on the device.

    movie = StokesMovie(data, times, nus, sigma, mask)          # data: (n_alpha, n_beta, nt, nnu, 4)
    chi   = chi2(params, movie, cache, L)
    chi   = chi2_gradient(dparams, params, movie, cache, L)     # dparams += d chi2 / d params
    fwd, rev = sweep_passes(dparams, cache, params, L)
    fs    = FrameScans(u, v, vis, sigma)
    chi_f = frame_chi2_seed(cache, c, fs, dalpha, L, D)         # seed of frame c left on the device

The seeds stay in HBM between the χ² and the reverse pass; only the χ² scalars and the gradient
come back to the host.
"""
import numpy as np

from . import api
from .api import DimensionMismatch, Model


class StokesMovie:
    """StokesMovie(data, times, νs, σ; mask) (Fit.jl:24-41): observed Stokes cube
    (n_alpha, n_beta, nt, nnu, 4) [cgs], observation times [M], frequencies [Hz], the noise per
    Stokes parameter (a 4-vector or an array shaped like data) and a boolean mask
    (n_alpha, n_beta, nt, nnu) of the pixels that enter χ²."""

    def __init__(self, data, times, nus, sigma, mask=None):
        self.data = np.asarray(data, dtype=np.float64)
        if self.data.ndim != 5 or self.data.shape[-1] != 4:
            raise DimensionMismatch(f"data must be (n_alpha, n_beta, nt, nnu, 4), got {self.data.shape}")
        self.times = np.asarray(times, dtype=np.float64).ravel()
        self.nus = np.asarray(nus, dtype=np.float64).ravel()
        if self.data.shape[2:4] != (len(self.times), len(self.nus)):
            raise DimensionMismatch("data's frame and frequency axes must match times and nus")
        self.sigma = np.asarray(sigma, dtype=np.float64)
        if self.sigma.shape != (4,) and self.sigma.shape != self.data.shape:
            raise DimensionMismatch("sigma must be a 4-vector or shaped like data")
        self.mask = np.ones(self.data.shape[:4], dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
        if self.mask.shape != self.data.shape[:4]:
            raise DimensionMismatch("mask must be (n_alpha, n_beta, nt, nnu)")

    def noise(self, k, l):
        return self.sigma if self.sigma.shape == (4,) else self.sigma[:, :, k, l, :]


def _frames(movie, frames, freqs):
    ks = list(range(len(movie.times))) if frames is None else [int(k) for k in frames]
    ls = list(range(len(movie.nus))) if freqs is None else [int(l) for l in freqs]
    return ks, ls


def chi2(params, movie, cache, L, frames=None, freqs=None, nmax=-1, slab=0.0, binning=None):
    """chi2(params, movie, cache, L; frames, freqs, nmax, slab, binning) (Fit.jl:66-93): Σ ((model -
    data)/σ)² over the masked pixels of the selected frames and frequencies, the model integrated over
    the pixels of a Binning when one is given (pixel_stokes, Fit.jl:73-75); the model cube is rendered
    in one launch on the device (the cache's precision policy applies), the sum is taken on the host."""
    ks, ls = _frames(movie, frames, freqs)
    cube = api.polarized_cube(cache, params, movie.times[ks], movie.nus[ls], L, nmax=nmax, slab=slab)
    if binning is not None:
        cube = api.bin(binning, cube)
    total = 0.0
    for b, l in enumerate(ls):
        for a, k in enumerate(ks):
            r = (cube[:, :, a, b, :] - movie.data[:, :, k, l, :]) / movie.noise(k, l)
            total += float(np.sum((r * r)[movie.mask[:, :, k, l]]))
    return total


def sweep_passes(dparams, cache, params, L, nmax=-1, slab=0.0):
    """sweep_passes(dparams, cache, params, L; method = :dual, nmax, slab) (Fit.jl:451-474): the two
    passes of the device gradient as closures. forward(t, nu, device=False) returns the model images
    in sorted order ((npix, 4) for a scalar time, (nframes, npix, 4) for a vector) and keeps the
    records; reverse(dstokes, t, nu) adds d(dstokes . image)/dparams into dparams, dstokes in sorted
    order or None for the seeds a χ² seed call left on the device."""
    m = params if isinstance(params, Model) else Model(params)

    def forward(t, nu, device=False):
        ts = np.atleast_1d(np.asarray(t, dtype=np.float64))
        img = api.gradient_forward(cache, m, ts, [nu], L, nmax=nmax, slab=slab, copy=not device)
        if device:
            return None
        return img[0] if np.ndim(t) else img[0, 0]

    def reverse(dstokes, t, nu):
        ts = np.atleast_1d(t)
        ds = None if dstokes is None else np.asarray(dstokes, dtype=np.float64).reshape(1, len(ts), cache.npix, 4)
        api.gradient_reverse(cache, m, dparams, ds)

    return forward, reverse


def chi2_gradient(dparams, params, movie, cache, L, frames=None, freqs=None, nmax=-1, slab=0.0, batch_frames=4, binning=None):
    """chi2_gradient!(dparams, params, movie, cache, L; frames, freqs, nmax, slab, batch_frames, binning)
    (Fit.jl:476-534): the movie χ² and, added into dparams, its gradient with respect to the thermal
    splat parameters. Per frequency and batch of frames: one forward launch, the image χ² and its
    seeds 2 r / σ² on the device (_frame_seed!, Fit.jl:537-577, through the Binning when given), one
    reverse launch."""
    m = params if isinstance(params, Model) else Model(params)
    ks, ls = _frames(movie, frames, freqs)
    total = 0.0
    for l in ls:
        for c0 in range(0, len(ks), max(int(batch_frames), 1)):
            chunk = ks[c0:c0 + max(int(batch_frames), 1)]
            api.gradient_forward(cache, m, movie.times[chunk], [movie.nus[l]], L, nmax=nmax, slab=slab, copy=False)
            for c, k in enumerate(chunk):
                total += api.image_chi2_seed(cache, c, movie.data[:, :, k, l, :], movie.noise(k, l), movie.mask[:, :, k, l], binning=binning)
            api.gradient_reverse(cache, m, dparams, None)
    return total


class FrameScans:
    """FrameScans(u, v, vis, σ, tap, g) (device_visibilities.jl:13-25): the baselines of one frame
    [wavelengths], the observed Stokes visibilities (complex (nuv, 4) [Jy]), their noise
    ((nuv, 4) or a 4-vector [Jy]), the taper per baseline (None = 1) and the gain factor
    g_{s1} conj(g_{s2}) per baseline (None = 1)."""

    def __init__(self, u, v, vis, sigma, tap=None, gains=None):
        self.u = np.asarray(u, dtype=np.float64).ravel()
        self.v = np.asarray(v, dtype=np.float64).ravel()
        self.vis = np.asarray(vis, dtype=np.complex128)
        self.sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (len(self.u), 4)).copy()
        self.tap = None if tap is None else np.asarray(tap, dtype=np.float64).ravel()
        self.gains = None if gains is None else np.asarray(gains, dtype=np.complex128).ravel()


def frame_chi2_seed(cache, c, fs, dalpha, L, D, return_model=False):
    """frame_chi2_seed!(seed, c, images, fs, cache, Δα, L, D) (device_visibilities.jl:97-111): the
    χ² of frame c of the last forward pass against its scans, with the adjoint seed of the weighted
    residuals written on the device for the reverse pass (reverse(None, t, nu))."""
    psize = dalpha * L / D
    return api.vis_chi2_seed(cache, c, psize, fs.u, fs.v, fs.vis, fs.sigma, taper=fs.tap, gains=fs.gains, return_model=return_model)


def image_loss_gradient(dparams, loss_and_grad, cache, params, t_obs, nu, L, nmax=-1, slab=0.0, binning=None):
    """image_loss_gradient!(dparams, loss, cache, params, t_obs, ν, L; binning) (Fit.jl:586-615) with the
    image derivative supplied by the caller: loss_and_grad(image) takes the screen-shaped observed
    image (n_alpha, n_beta, 4), binned over the pixels of a Binning when one is given, and returns
    (value, d value / d image); the reverse pass on the device turns that seed into d loss / d params,
    added into dparams. Returns the loss value."""
    m = params if isinstance(params, Model) else Model(params)
    sorted_img = api.gradient_forward(cache, m, [t_obs], [nu], L, nmax=nmax, slab=slab)[0, 0]
    na, nb = cache.screen_size
    perm = cache.perm_host - 1                      # screen index of every sorted pixel
    screen = np.zeros((cache.npix, 4))
    screen[perm] = sorted_img
    if binning is None:
        value, dimg = loss_and_grad(np.transpose(screen.reshape(nb, na, 4), (1, 0, 2)))
        dflat = np.transpose(np.asarray(dimg, dtype=np.float64), (1, 0, 2)).reshape(cache.npix, 4)
    else:
        value, dimg = loss_and_grad(api.bin(binning, screen))
        bna, bnb = binning.size
        dpix = np.transpose(np.asarray(dimg, dtype=np.float64), (1, 0, 2)).reshape(bna * bnb, 4)
        dflat = dpix[binning.pixel] / binning.count[binning.pixel][:, None]      # the mean's adjoint
    api.gradient_reverse(cache, m, dparams, dflat[perm][None, None])
    return value


# ---- closure quantities and station gains (visibilities.jl:141-214) --------------------------------------
def _signed_vis(vis, k):
    """Stokes I visibility of 1-based row k, conjugated for a negative k (the stored baseline runs the other way)."""
    return np.conj(vis[-k - 1, 0]) if k < 0 else vis[k - 1, 0]


def closure_phases(vis, triangles):
    """closure_phases(vis, triangles) (visibilities.jl:141-150): arg(V1 V2 V3) of Stokes I for triangles of
    1-based visibility indices closing (ij, jk, ki); a negative index conjugates. vis: complex (nuv, 4)."""
    vis = np.asarray(vis)
    return np.array([np.angle(_signed_vis(vis, t[0]) * _signed_vis(vis, t[1]) * _signed_vis(vis, t[2])) for t in triangles])


def log_closure_amplitudes(vis, quadrangles):
    """log_closure_amplitudes(vis, quadrangles) (visibilities.jl:152-161): ln(|V12||V34| / (|V13||V24|)) of
    Stokes I for quadrangles of 1-based indices (12, 34, 13, 24); signs are ignored."""
    vis = np.asarray(vis)
    a = lambda k: np.abs(vis[abs(k) - 1, 0])
    return np.array([np.log(a(q[0]) * a(q[1]) / (a(q[2]) * a(q[3]))) for q in quadrangles])


class ClosureData:
    """ClosureData(u, v, triangles, phases, σ_phase, quadrangles, logamps, σ_logamp) (visibilities.jl:163-178)."""

    def __init__(self, u, v, triangles=(), phases=(), sigma_phase=(), quadrangles=(), logamps=(), sigma_logamp=()):
        self.u = np.asarray(u, dtype=np.float64).ravel()
        self.v = np.asarray(v, dtype=np.float64).ravel()
        self.triangles = [tuple(int(x) for x in t) for t in triangles]
        self.phases = np.asarray(phases, dtype=np.float64)
        self.sigma_phase = np.asarray(sigma_phase, dtype=np.float64)
        self.quadrangles = [tuple(int(x) for x in q) for q in quadrangles]
        self.logamps = np.asarray(logamps, dtype=np.float64)
        self.sigma_logamp = np.asarray(sigma_logamp, dtype=np.float64)


def closure_chi2(V, data, grad=False):
    """The closure χ² of model visibilities V (complex (nuv, 4)) as chi2_closures computes it
    (visibilities.jl:180-199: closure-phase differences wrapped to (-pi, pi]); with grad=True also
    w = dchi2/dRe V + i dchi2/dIm V, the input of vis_seed (Stokes I only)."""
    V = np.asarray(V, dtype=np.complex128)
    total = 0.0
    w = np.zeros_like(V)
    if data.triangles:
        cp = closure_phases(V, data.triangles)
        d = np.remainder(cp - data.phases + np.pi, 2 * np.pi) - np.pi
        total += float(np.sum((d / data.sigma_phase) ** 2))
        if grad:
            coef = 2.0 * d / data.sigma_phase ** 2
            for c, t in zip(coef, data.triangles):
                for k in t:      # d arg(V) = Im(dV / V): w = i V / |V|^2, negated for a conjugated leg
                    z = V[abs(k) - 1, 0]
                    w[abs(k) - 1, 0] += c * (1j if k > 0 else -1j) * z / abs(z) ** 2
    if data.quadrangles:
        la = log_closure_amplitudes(V, data.quadrangles)
        r = (la - data.logamps) / data.sigma_logamp
        total += float(np.sum(r ** 2))
        if grad:
            coef = 2.0 * r / data.sigma_logamp
            for c, q in zip(coef, data.quadrangles):
                for sgn, k in zip((1.0, 1.0, -1.0, -1.0), q):     # d ln|V| = Re(dV / V): w = V / |V|^2
                    z = V[abs(k) - 1, 0]
                    w[abs(k) - 1, 0] += c * sgn * z / abs(z) ** 2
    return (total, w) if grad else total


def frame_chi2_closures(cache, c, data, dalpha, L, D, taper=None, add=False, on_host=False):
    """chi2_closures of frame c of the last forward pass with its adjoint seed written on the device (added to the
    frame's seed with add=True: a visibility and a closure term in one likelihood): on the device
    (api.closure_chi2_seed), or with on_host=True the visibilities copied to the host, the closures and their
    gradient there (closure_chi2) and the seed through vis_seed, as the reference computes them."""
    psize = dalpha * L / D
    if not on_host:
        return api.closure_chi2_seed(cache, c, psize, data, taper=taper, add=add)
    V = api.visibilities(cache, c + 1, psize, data.u, data.v, taper=taper)[c]
    chi, w = closure_chi2(V, data, grad=True)
    api.vis_seed(cache, c, psize, data.u, data.v, w, taper=taper, add=add)
    return chi


def gain_factors(gains, s1, s2):
    """g_{s1} conj(g_{s2}) per baseline from a 2 x nstations gain matrix (log-amplitude, phase) and 1-based
    station indices (visibilities.jl:92)."""
    g = np.asarray(gains, dtype=np.float64)
    s1, s2 = np.asarray(s1) - 1, np.asarray(s2) - 1
    return np.exp(g[0, s1] + g[0, s2]) * np.exp(1j * (g[1, s1] - g[1, s2]))


def apply_gains(vis, gains, s1, s2):
    """apply_gains(vis, gains, s1, s2) (visibilities.jl:201-214): V'_k = g_{s1[k]} conj(g_{s2[k]}) V_k."""
    return np.asarray(vis) * gain_factors(gains, s1, s2)[:, None]


# ---- the spacetime block (joint.jl:74-154) ---------------------------------------------------------------
def _spacetime_tangent(x, camera, N, params, times, nus, nmax, slab, kind, stored):
    """The cube of `camera` at x = (a, theta_o, ln L) with N samples per ray and its partials in x, on the device
    (Richardson-extrapolated central differences of FP64 cubes): stored Recurrence(64) samples, or the fused march
    (joint.jl:78)."""
    return api.spacetime_tangent_differences(camera, N, float(x[0]), float(x[1]), params, times, nus, float(np.exp(x[2])),
                                             nmax=nmax, slab=slab, kind=kind, marcher=None if stored else api.Fused(64))


def spacetime_valgrad(loss_and_grad, x, camera, N, params, times, nus, nmax=-1, slab=0.0, kind="thermal", stored=True):
    """spacetime_valgrad(loss, x, camera, backend; N) (joint.jl:74-82) for a loss of the rendered cube: the
    value and gradient with respect to x = (a, theta_o, ln L). The cube of `params` at `times`, `nus` and its
    partials come from spacetime_tangent_differences; loss_and_grad(cube) takes the screen-shaped cube
    (n_alpha, n_beta, nt, nnu, 4) and returns (value, d value / d cube) (spacetime_movie_loss builds the
    movie χ²). Returns (value, gradient (3,))."""
    cube, dcube = _spacetime_tangent(x, camera, N, params, times, nus, nmax, slab, kind, stored)
    value, dv = loss_and_grad(cube)
    g = np.einsum("abtfs,abtfsk->k", np.asarray(dv, dtype=np.float64), dcube)
    return float(value), g


def spacetime_movie_loss(movie, frames=None, freqs=None, binning=None):
    """The movie χ² as a loss of the rendered cube for spacetime_valgrad (spacetime_movie_loss, joint.jl:90-96):
    the cube must cover the movie's selected times and frequencies in order; with a Binning the cube is
    integrated over its pixels first. Returns loss_and_grad(cube) -> (χ², dχ²/dcube)."""
    ks, ls = _frames(movie, frames, freqs)

    def loss_and_grad(cube):
        model = api.bin(binning, cube) if binning is not None else cube
        total = 0.0
        dmodel = np.zeros(model.shape)
        for b, l in enumerate(ls):
            for a, k in enumerate(ks):
                sg = movie.noise(k, l)
                r = (model[:, :, a, b, :] - movie.data[:, :, k, l, :]) / sg
                mk = movie.mask[:, :, k, l][..., None]
                total += float(np.sum((r * r) * mk))
                dmodel[:, :, a, b, :] = np.where(mk, 2 * r / sg, 0.0)
        if binning is None:
            return total, dmodel
        flat = np.transpose(dmodel, (1, 0) + tuple(range(2, dmodel.ndim))).reshape((binning.npixels,) + dmodel.shape[2:])
        pts = flat[binning.pixel] / binning.count[binning.pixel].reshape((-1,) + (1,) * (flat.ndim - 1))
        return total, pts.reshape((binning.npoints, 1) + dmodel.shape[2:])       # the cube of a point camera

    return loss_and_grad


def spacetime_jacobian(x, camera, N, params, movie, frames=None, freqs=None, nmax=-1, slab=0.0, binning=None, kind="thermal",
                       stored=True):
    """spacetime_jacobian with the movie residuals (joint.jl:98-154, spacetime_movie_residuals): the scaled
    residuals (model - data)/σ of the masked pixels, ordered frequency, frame, β, α, Stokes as the reference
    appends them, and their Jacobian with respect to x = (a, theta_o, ln L). Rows with a finite residual and
    non-finite partials are zeroed, with a warning above one per thousand. Returns (r, J)."""
    if len(x) != 3:
        raise ValueError("the spacetime block is (a, theta_o, ln L)")
    ks, ls = _frames(movie, frames, freqs)
    cube, dcube = _spacetime_tangent(x, camera, N, params, movie.times[ks], movie.nus[ls], nmax, slab, kind, stored)
    if binning is not None:
        cube = api.bin(binning, cube)
        dcube = api.bin(binning, dcube)
    rs, Js = [], []
    for b, l in enumerate(ls):
        for a, k in enumerate(ks):
            sg = np.broadcast_to(movie.noise(k, l), movie.data.shape[:2] + (4,))
            r = (cube[:, :, a, b, :] - movie.data[:, :, k, l, :]) / sg
            J = dcube[:, :, a, b, :, :] / sg[..., None]
            mk = movie.mask[:, :, k, l]
            rr = np.transpose(r, (1, 0, 2))[mk.T]                     # j (beta) outer, i (alpha) inner
            JJ = np.transpose(J, (1, 0, 2, 3))[mk.T]
            rs.append(rr.reshape(-1))
            Js.append(JJ.reshape(-1, 3))
    r = np.concatenate(rs) if rs else np.zeros(0)
    J = np.concatenate(Js) if Js else np.zeros((0, 3))
    bad = np.isfinite(r) & ~np.all(np.isfinite(J), axis=1)
    if bad.any():
        if bad.sum() > 1e-3 * len(r):
            import warnings
            warnings.warn(f"spacetime Jacobian: {int(bad.sum())} of {len(r)} residuals have non-finite partials; their rows are zeroed")
        J[bad] = 0.0
    return r, J


# ---- the diffractive scattering kernel of Sgr A* (scattering.jl:1-38) ------------------------------------
class ScatteringKernel:
    """ScatteringKernel(ν; fwhm_maj = 1.380, fwhm_min = 0.703, pa = 81.9) (scattering.jl:5-34): the
    ensemble-average scattering of Sgr A* at frequency ν [Hz] as a Gaussian in the visibility domain, FWHM
    fwhm_maj mas λ_cm² along the position angle pa (degrees east of north) and fwhm_min mas λ_cm² across it
    (Johnson et al. 2018; ehtim's Bower et al. 2006 constants are 1.309, 0.640, 78). kernel(u, v) at
    (u, v) in wavelengths is exp(-2π² (a u² + 2 c u v + b v²)), ehtim's covariance form; it is the taper
    per baseline that FrameScans and the device χ² take."""

    def __init__(self, nu, fwhm_maj=1.380, fwhm_min=0.703, pa=81.9):
        lam_cm = 2.99792458e10 / nu
        mas = 1e-3 * np.pi / (180 * 3600)
        to_sigma = lam_cm ** 2 * mas / (2 * np.sqrt(2 * np.log(2)))
        smaj, smin = fwhm_maj * to_sigma, fwhm_min * to_sigma
        th = -np.deg2rad(pa)                      # the angle enters negated in ehtim's convention
        ct, st = np.cos(th), np.sin(th)
        self.a = (smin * ct) ** 2 + (smaj * st) ** 2
        self.b = (smaj * ct) ** 2 + (smin * st) ** 2
        self.c = (smin ** 2 - smaj ** 2) * ct * st

    def __call__(self, u, v):
        u, v = np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64)
        return np.exp(-2 * np.pi ** 2 * (self.a * u * u + 2 * self.c * u * v + self.b * v * v))


def taper(kernel, model, u, v):
    """taper(kernel, model, u, v) (scattering.jl:36-38): model visibilities (nuv, 4) times the kernel at
    each baseline; kernel None leaves them unchanged."""
    if kernel is None:
        return model
    return np.asarray(model) * kernel(u, v)[:, None]


# ---- the fitting schedule: stages, Adam, annealing, partition hygiene (Fit.jl:94-275, 308-438, 620-659) -------
class Stage:
    """Stage(; free, iterations, eta, eta_end, batch, freqs, steps, label) (Fit.jl:318-338): the rows to free (0-based
    indices; None frees all), Adam iterations, the learning rate annealed from eta to eta_end on a cosine, the
    minibatch (nframes, nfreqs) or None, the frequency indices that enter (None = all) and per-row step
    multipliers {row: factor}."""

    def __init__(self, free=None, iterations=100, eta=0.02, eta_end=None, batch=None, freqs=None, steps=None, label=""):
        self.free, self.iterations, self.eta = free, int(iterations), float(eta)
        self.eta_end = self.eta if eta_end is None else float(eta_end)
        self.batch, self.freqs, self.steps, self.label = batch, freqs, steps, label


class Hygiene:
    """Hygiene(; every, prune_fraction, densify_threshold, merge_position, merge_shape, max_splats) (Fit.jl:340-356):
    prune, merge and densify every `every` iterations of a stage (0 disables); each pass resets the optimizer."""

    def __init__(self, every=0, prune_fraction=1e-4, densify_threshold=np.inf, merge_position=0.2, merge_shape=0.1, max_splats=2 ** 62):
        self.every, self.prune_fraction, self.densify_threshold = int(every), float(prune_fraction), float(densify_threshold)
        self.merge_position, self.merge_shape, self.max_splats = float(merge_position), float(merge_shape), int(max_splats)


def freeze(params, rows):
    """freeze(params, rows) (Fit.jl:96-103): a boolean mask that frees the given rows (0-based) of every splat."""
    free = np.zeros(np.shape(params), dtype=bool)
    for r in rows:
        if not 0 <= int(r) < free.shape[0]:
            raise ValueError(f"unknown parameter row {r}")
        free[int(r), :] = True
    return free


def step_scale(params, steps):
    """step_scale(params, steps) (Fit.jl:117-126): per-entry multipliers of the Adam step, 1 except for the rows
    given in steps ({row: factor}); applied as old + scale (new - old)."""
    scale = np.ones(np.shape(params))
    if steps is None:
        return scale
    for r, f in dict(steps).items():
        if not 0 <= int(r) < scale.shape[0]:
            raise ValueError(f"unknown parameter row {r}")
        scale[int(r), :] = f
    return scale


class Adam:
    """Adam (Kingma & Ba 2015) with Optimisers.jl's defaults, beta = (0.9, 0.999), epsilon = 1e-8: per entry,
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, x -= eta (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)."""

    def __init__(self, shape, eta, beta=(0.9, 0.999), eps=1e-8):
        self.eta, self.b1, self.b2, self.eps = float(eta), beta[0], beta[1], eps
        self.m, self.v = np.zeros(shape), np.zeros(shape)
        self.p1, self.p2 = 1.0, 1.0

    def step(self, x, g):
        self.p1 *= self.b1
        self.p2 *= self.b2
        self.m = self.b1 * self.m + (1.0 - self.b1) * g
        self.v = self.b2 * self.v + (1.0 - self.b2) * g * g
        return x - self.eta * (self.m / (1.0 - self.p1)) / (np.sqrt(self.v / (1.0 - self.p2)) + self.eps)


def _rot(q):
    w, x, y, z = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def prune(params, fraction=1e-4):
    """prune(params; fraction) (Fit.jl:172-176): drop splats whose density x volume e^(ln n_e + s1 + s2 + s3) is
    below fraction of the largest; returns (params, kept indices)."""
    P = np.asarray(params)
    w = np.exp(P[12] + P[3] + P[4] + P[5])
    kept = np.nonzero(w >= fraction * w.max())[0]
    return np.asfortranarray(P[:, kept]), kept


def densify(params, grad, threshold, scale_factor=1.6):
    """densify(params, grad; threshold, scale_factor) (Fit.jl:187-209): split every splat whose centre-gradient
    norm exceeds threshold into two at +-1/2 of its largest principal scale, every scale divided by scale_factor
    and the density set so that the pair's density x volume equals the parent's."""
    P, G = np.asarray(params), np.asarray(grad)
    cols = []
    for i in range(P.shape[1]):
        col = P[:, i]
        if np.sqrt(G[0, i] ** 2 + G[1, i] ** 2 + G[2, i] ** 2) > threshold:
            k = int(np.argmax(col[3:6]))
            axis = _rot(col[6:10])[:, k] * np.exp(col[3 + k]) / 2
            for sgn in (-1.0, 1.0):
                c = col.copy()
                c[0:3] += sgn * axis
                c[3:6] -= np.log(scale_factor)
                c[12] += 3 * np.log(scale_factor) - np.log(2.0)
                cols.append(c)
        else:
            cols.append(col)
    return np.asfortranarray(np.stack(cols, axis=1))


def merge(params, position_tol=0.2, shape_tol=0.1):
    """merge(params; position_tol, shape_tol) (Fit.jl:217-240): splats whose centres lie within position_tol and
    whose log scales agree within shape_tol become one with the summed density and the density-weighted mean of
    the other rows; returns (params, groups), greedily in index order."""
    P = np.asarray(params)
    n = P.shape[1]
    assigned = np.zeros(n, dtype=int)
    groups = []
    for i in range(n):
        if assigned[i]:
            continue
        g = [i]
        assigned[i] = len(groups) + 1
        for j in range(i + 1, n):
            if assigned[j]:
                continue
            if np.sqrt(np.sum((P[0:3, i] - P[0:3, j]) ** 2)) <= position_tol and np.max(np.abs(P[3:6, i] - P[3:6, j])) <= shape_tol:
                g.append(j)
                assigned[j] = assigned[i]
        groups.append(g)
    out = np.empty((P.shape[0], len(groups)))
    for k, g in enumerate(groups):
        w = np.exp(P[12, g])
        out[:, k] = (P[:, g] * (w / w.sum())).sum(axis=1)
        out[12, k] = np.log(w.sum())
    return np.asfortranarray(out), groups


def fit_loop(params, make_valgrad, stages, hygiene=None, callback=None):
    """_fit_loop!(params, make_valgrad, stages; hygiene, callback) (Fit.jl:620-659): for each stage, Adam on the
    free rows with the cosine-annealed rate; make_valgrad(stage, it) returns a function q -> (chi2, gradient);
    every hygiene.every iterations prune, merge and densify (the splat count changes; the optimizer restarts).
    Returns (params, history, events)."""
    hygiene = hygiene or Hygiene()
    params = np.asfortranarray(np.array(params, dtype=np.float64))
    history, events = [], []
    for si, st in enumerate(stages, start=1):
        mask = np.ones(params.shape) if st.free is None else freeze(params, st.free).astype(np.float64)
        scale = None if st.steps is None else step_scale(params, st.steps)
        opt = Adam(params.shape, st.eta)
        for it in range(1, st.iterations + 1):
            eta = st.eta_end + (st.eta - st.eta_end) * (1 + np.cos(np.pi * (it - 1) / max(st.iterations - 1, 1))) / 2
            opt.eta = eta
            value, g = make_valgrad(st, it)(params)
            g = np.asarray(g)
            if hygiene.every > 0 and it % hygiene.every == 0:
                before = params.shape[1]
                q, kept = prune(params, hygiene.prune_fraction)
                gk = g[:, kept]
                q, groups = merge(q, hygiene.merge_position, hygiene.merge_shape)
                gm = np.stack([gk[:, grp].sum(axis=1) for grp in groups], axis=1)
                if q.shape[1] < hygiene.max_splats:
                    q = densify(q, gm, hygiene.densify_threshold)
                if q.shape != params.shape or not np.array_equal(q, params):
                    params = q
                    events.append((si, it, before, params.shape[1]))
                    mask = np.ones(params.shape) if st.free is None else freeze(params, st.free).astype(np.float64)
                    scale = None if st.steps is None else step_scale(params, st.steps)
                    opt = Adam(params.shape, eta)
                    continue
            old = params
            new = opt.step(params, g * mask)
            params = np.asfortranarray(new if scale is None else old + scale * (new - old))
            history.append(value)
            if callback is not None:
                callback(si, it, params, value)
    return params, history, events


def fit(params, movie, cache, L, stages, hygiene=None, rng=None, callback=None, nmax=-1, slab=0.0, binning=None, kind="thermal",
        graph=False):
    """fit!(params, movie, cache, L, stages; hygiene, rng, callback, nmax, slab, gradient = :dual, binning)
    (Fit.jl:366-404): the staged Adam schedule on the movie chi2 with the gradient of the device dual sweep
    (chi2_gradient); minibatches of frames and frequencies drawn per iteration when a stage asks for them. The
    parameters live on the host (hygiene changes their count) and a device model is built from them every
    iteration. With graph, iterations over the whole movie (no minibatch, no binning, no half-orbit cut) replay
    one captured graph (api.FitGraph) with the new parameters, captured again when the splat count changes.
    Returns (params, history, events)."""
    rng = rng or np.random.default_rng()
    fg = {}
    if graph:
        if binning is not None or nmax >= 0:
            raise ValueError("the captured fit iteration takes the plain image chi2 (no binning, no half-orbit cut)")
        fg["graph"] = api.FitGraph(cache, movie.times, movie.nus, L, image=(movie.data, movie.sigma, movie.mask))

    def batches(st):
        freqs_all = list(range(len(movie.nus))) if st.freqs is None else list(st.freqs)
        if st.batch is None:
            return list(range(len(movie.times))), freqs_all
        frames = sorted(rng.permutation(len(movie.times))[:min(st.batch[0], len(movie.times))])
        freqs = sorted(np.asarray(freqs_all)[rng.permutation(len(freqs_all))[:min(st.batch[1], len(freqs_all))]])
        return frames, freqs

    def make_valgrad(st, it):
        if graph and (st.batch is not None or st.freqs is not None):
            raise ValueError("the captured fit iteration covers the whole movie: stages without minibatches")
        frames, freqs = batches(st)

        def valgrad(q):
            if graph:
                m = fg.get("model")
                if m is None or m.ns != q.shape[1]:
                    m = fg["model"] = Model(q, kind=kind)
                return fg["graph"].step(m, q)
            m = Model(q, kind=kind)
            g = np.zeros(q.shape, order="F")
            chi = chi2_gradient(g, m, movie, cache, L, frames=frames, freqs=freqs, nmax=nmax, slab=slab, binning=binning)
            return chi, g
        return valgrad

    return fit_loop(params, make_valgrad, stages, hygiene=hygiene, callback=callback)
