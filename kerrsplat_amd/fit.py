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


def chi2(params, movie, cache, L, frames=None, freqs=None, nmax=-1, slab=0.0):
    """chi2(params, movie, cache, L; frames, freqs, nmax, slab) (Fit.jl:66-93): Σ ((model - data)/σ)²
    over the masked pixels of the selected frames and frequencies; the model cube is rendered in one
    launch on the device (the cache's precision policy applies), the sum is taken on the host."""
    ks, ls = _frames(movie, frames, freqs)
    cube = api.polarized_cube(cache, params, movie.times[ks], movie.nus[ls], L, nmax=nmax, slab=slab)
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


def chi2_gradient(dparams, params, movie, cache, L, frames=None, freqs=None, nmax=-1, slab=0.0, batch_frames=4):
    """chi2_gradient!(dparams, params, movie, cache, L; frames, freqs, nmax, slab, batch_frames)
    (Fit.jl:476-534): the movie χ² and, added into dparams, its gradient with respect to the thermal
    splat parameters. Per frequency and batch of frames: one forward launch, the image χ² and its
    seeds 2 r / σ² on the device (_frame_seed!, Fit.jl:537-577), one reverse launch."""
    m = params if isinstance(params, Model) else Model(params)
    ks, ls = _frames(movie, frames, freqs)
    total = 0.0
    for l in ls:
        for c0 in range(0, len(ks), max(int(batch_frames), 1)):
            chunk = ks[c0:c0 + max(int(batch_frames), 1)]
            api.gradient_forward(cache, m, movie.times[chunk], [movie.nus[l]], L, nmax=nmax, slab=slab, copy=False)
            for c, k in enumerate(chunk):
                total += api.image_chi2_seed(cache, c, movie.data[:, :, k, l, :], movie.noise(k, l), movie.mask[:, :, k, l])
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


def image_loss_gradient(dparams, loss_and_grad, cache, params, t_obs, nu, L, nmax=-1, slab=0.0):
    """image_loss_gradient!(dparams, loss, cache, params, t_obs, ν, L) (Fit.jl:586-615) with the
    image derivative supplied by the caller: loss_and_grad(image) takes the screen-shaped observed
    image (n_alpha, n_beta, 4) and returns (value, d value / d image); the reverse pass on the device
    turns that seed into d loss / d params, added into dparams. Returns the loss value."""
    m = params if isinstance(params, Model) else Model(params)
    sorted_img = api.gradient_forward(cache, m, [t_obs], [nu], L, nmax=nmax, slab=slab)[0, 0]
    na, nb = cache.screen_size
    perm = cache.perm_host - 1                      # screen index of every sorted pixel
    screen = np.zeros((cache.npix, 4))
    screen[perm] = sorted_img
    value, dimg = loss_and_grad(np.transpose(screen.reshape(nb, na, 4), (1, 0, 2)))
    dflat = np.transpose(np.asarray(dimg, dtype=np.float64), (1, 0, 2)).reshape(cache.npix, 4)
    api.gradient_reverse(cache, m, dparams, dflat[perm][None, None])
    return value
