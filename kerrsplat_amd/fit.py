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


def frame_chi2_closures(cache, c, data, dalpha, L, D, taper=None, add=False):
    """chi2_closures of frame c of the last forward pass (visibilities on the device, closures on the host,
    as the reference computes them) with its adjoint seed written on the device through vis_seed
    (added to the frame's seed with add=True: a visibility and a closure term in one likelihood)."""
    psize = dalpha * L / D
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
def _spacetime_cache(x, camera, N):
    """A cache of `camera` with N samples per ray at (a, theta_o) = x[:2] (pixel constants only: the tangent
    kernel marches the closed-form samples itself)."""
    cache = api.GeodesicCache(camera, N, store_samples=False)
    api.regenerate(cache, float(x[0]), float(x[1]), marcher=api.Fused(64))
    return cache


def spacetime_valgrad(loss_and_grad, x, camera, N, params, times, nus, nmax=-1, slab=0.0, kind="thermal"):
    """spacetime_valgrad(loss, x, camera, backend; N) (joint.jl:74-82) for a loss of the rendered cube: the
    value and gradient with respect to x = (a, theta_o, ln L). The cube of `params` at `times`, `nus` and its
    forward-mode partials come from spacetime_tangent; loss_and_grad(cube) takes the screen-shaped cube
    (n_alpha, n_beta, nt, nnu, 4) and returns (value, d value / d cube) (spacetime_movie_loss builds the
    movie χ²). Returns (value, gradient (3,))."""
    cache = _spacetime_cache(x, camera, N)
    cube, dcube = api.spacetime_tangent(cache, params, times, nus, float(np.exp(x[2])), nmax=nmax, slab=slab, kind=kind)
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


def spacetime_jacobian(x, camera, N, params, movie, frames=None, freqs=None, nmax=-1, slab=0.0, binning=None, kind="thermal"):
    """spacetime_jacobian with the movie residuals (joint.jl:98-154, spacetime_movie_residuals): the scaled
    residuals (model - data)/σ of the masked pixels, ordered frequency, frame, β, α, Stokes as the reference
    appends them, and their Jacobian with respect to x = (a, theta_o, ln L). Rows with a finite residual and
    non-finite partials are zeroed, with a warning above one per thousand. Returns (r, J)."""
    if len(x) != 3:
        raise ValueError("the spacetime block is (a, theta_o, ln L)")
    ks, ls = _frames(movie, frames, freqs)
    cache = _spacetime_cache(x, camera, N)
    cube, dcube = api.spacetime_tangent(cache, params, movie.times[ks], movie.nus[ls], float(np.exp(x[2])), nmax=nmax, slab=slab, kind=kind)
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
