"""Python mirror of the reference API over the C-ABI (include/kerrsplat_hip.h).

Names, argument meaning and errors follow the reference's Julia functions; each wrapper cites the
function it mirrors. Exceptions: ValueError for the reference's ArgumentError, IndexError-like
DimensionMismatch as `DimensionMismatch` (a ValueError subclass), RuntimeError (KernelError) for
HIP runtime failures.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libkerrsplat_hip.so"

KS_OK, KS_ERR_ARG, KS_ERR_DIM, KS_ERR_STATE, KS_ERR_HIP = 0, 1, 2, 3, 4
KS_MARCHER_DIRECT, KS_MARCHER_RECURRENCE, KS_MARCHER_FUSED = 0, 1, 2
MODEL_KIND = {"thin": 0, "thermal": 1, "powerlaw": 2, "kappa": 3, "knot": 4}
NSPLATPARAMS, NPOLARIZEDPARAMS, NPOWERLAWPARAMS, NKAPPAPARAMS = 14, 21, 22, 22
_NROWS = {"thin": 14, "thermal": 21, "knot": 21, "powerlaw": 22, "kappa": 22}

# cgs constants (src/Transfer/constants.jl:3-12)
GNEWT, MSUN, CL = 6.6742e-8, 1.989e33, 2.99792458e10


def gravitational_radius(m_solar):
    """L = GM/c^2 [cm] (transport.jl:16)."""
    return GNEWT * m_solar * MSUN / (CL * CL)


class KernelError(RuntimeError):
    pass


class DimensionMismatch(ValueError):
    pass


class StateError(ValueError):
    pass


def library_path():
    """The in-tree HIP library; KERRSPLAT_HIP_LIB names another build of it (to compare builds)."""
    return os.environ.get("KERRSPLAT_HIP_LIB") or os.path.join(_HERE, _LIB_NAME)


_lib = None
_ctx = None
_P = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise KernelError(f"{path} is missing: build it with __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    lib.ks_last_error.restype = ctypes.c_char_p
    lib.ks_ctx_stream.restype = _P
    lib.ks_stokes_device_ptr.restype = _P
    lib.ks_cache_create.argtypes = [_P, _i64, _i, _i, ctypes.POINTER(_P)]
    lib.ks_regenerate.argtypes = [_P, _P, _P, _i, _i, _d, _d, _i, _i]
    lib.ks_model_create.argtypes = [_P, _i, _i, _i, _P, _d, _P, _P, _i, ctypes.POINTER(_P)]
    lib.ks_model_composite.argtypes = [_P, _i, _P, ctypes.POINTER(_P)]
    lib.ks_render.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _d, _P, _i]
    lib.ks_spacetime_tangent.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _d, _P, _P]
    lib.ks_render_device.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _d]
    lib.ks_copy_stokes.argtypes = [_P, _i64, _P, _i]
    lib.ks_thin_image.argtypes = [_P, _P, _d, _P, _i]
    lib.ks_ray_lists.argtypes = [_P, _P, _P, _i, _i, _d, _i, _P, ctypes.POINTER(_i), _P]
    lib.ks_visibilities.argtypes = [_P, _i64, _d, _P, _P, _P, _i, _P]
    lib.ks_visibilities_noalloc.argtypes = [_P, _i64, _d, _P, _P, _P, _i, _P]
    lib.ks_timing_end.argtypes = [_P, ctypes.POINTER(ctypes.c_float)]
    lib.ks_stage_ms.argtypes = [_P, _i, ctypes.POINTER(ctypes.c_float)]
    lib.ks_polarized_gradient.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _d, _P, _i, _P, _P]
    lib.ks_gradient_forward.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _d, _P, _i]
    lib.ks_gradient_reverse.argtypes = [_P, _P, _P, _i, _P]
    lib.ks_vis_chi2_seed.argtypes = [_P, _i64, _d, _P, _P, _P, _P, _P, _P, _i, _P, _P]
    lib.ks_image_chi2_seed.argtypes = [_P, _i64, _P, _P, _i, _P, _P]
    lib.ks_binning_create.argtypes = [_P, _i64, _P, _i, _i, ctypes.POINTER(_P)]
    lib.ks_binning_destroy.argtypes = [_P]
    lib.ks_image_chi2_seed_binned.argtypes = [_P, _i64, _P, _P, _P, _i, _P, _P]
    lib.ks_vis_seed.argtypes = [_P, _i64, _d, _P, _P, _P, _P, _i, _i]
    lib.ks_closure_chi2_seed.argtypes = [_P, _i64, _d, _P, _P, _P, _i, _i, _P, _P, _P, _i, _P, _P, _P, _i, _P]
    lib.ks_march.argtypes = [_P]
    lib.ks_cache_precision.argtypes = [_P, _i]
    lib.ks_cache_march_precision.argtypes = [_P, _i]
    lib.ks_cache_tuning.argtypes = [_P, _i, _i, _i, _i]
    lib.ks_cache_get_tuning.argtypes = [_P, _P, _P, _P, _P]
    lib.ks_cache_set_parts.argtypes = [_P, _i, _P, _P]
    lib.ks_autotune.argtypes = [_P, _P, _P, _i, _P, _i, _d, _i, _i, _P, _i, ctypes.POINTER(_i)]
    lib.ks_queue_costs.argtypes = [_P, ctypes.POINTER(_i64), _P, _i]
    lib.ks_fit_graph_create.argtypes = [_P, _i, _P, _i, _P, _d, _i, _P, _P, _i, _P, _d, _P, _P, _P, _i, ctypes.POINTER(_P)]
    lib.ks_fit_graph_step.argtypes = [_P, _P, _P, _P, _P]
    lib.ks_fit_graph_info.argtypes = [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
    lib.ks_fit_graph_destroy.argtypes = [_P]
    _lib = lib
    return lib


KS_QUEUE_TRANSFER, KS_QUEUE_ADJOINT = 1, 2
KS_TUNE_RECORDS, KS_TUNE_ADJOINT, KS_TUNE_KEEP_TRANSFER = 1, 2, 4


def _check(rc):
    if rc == KS_OK:
        return
    msg = _load().ks_last_error().decode()
    if rc == KS_ERR_ARG:
        raise ValueError(msg)
    if rc == KS_ERR_DIM:
        raise DimensionMismatch(msg)
    if rc == KS_ERR_STATE:
        raise StateError(msg)
    raise KernelError(msg)


def context(device=0):
    """The process-wide device context (one GPU, one stream)."""
    global _ctx
    if _ctx is None:
        lib = _load()
        p = _P()
        _check(lib.ks_ctx_create(_i(device), ctypes.byref(p)))
        _ctx = p
    return _ctx


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- camera (src/Geodesics/camera.jl) ---------------------------------------------------------------------
def grid_axis(lo, hi, n):
    """Krang's screen formula lo + (hi - lo)(i - 1)/(n - 1) (camera.jl:122-124)."""
    if n == 1:
        return np.array([float(lo)])
    i = np.arange(1, n + 1, dtype=np.float64)
    return float(lo) + (float(hi) - float(lo)) * (i - 1) / (n - 1)


class Camera:
    """Camera((alo, ahi), (blo, bhi), res | (na, nb)) or Camera(alphas, betas) (camera.jl:97-120)."""

    def __init__(self, arange, brange=None, size=None):
        if brange is None or (size is None and np.ndim(arange) == 1 and len(arange) != 2):
            raise ValueError("Camera((alo, ahi), (blo, bhi), res) or Camera.points(alphas, betas)")
        if np.ndim(size) == 0:
            size = (int(size), int(size))
        na, nb = size
        av, bv = grid_axis(arange[0], arange[1], na), grid_axis(brange[0], brange[1], nb)
        self.alphas = np.repeat(av[:, None], nb, axis=1).ravel(order="F")
        self.betas = np.repeat(bv[None, :], na, axis=0).ravel(order="F")
        self.size = (na, nb)

    @classmethod
    def points(cls, alphas, betas, size=None):
        c = cls.__new__(cls)
        c.alphas = _f64(alphas).ravel()
        c.betas = _f64(betas).ravel()
        if len(c.alphas) != len(c.betas):
            raise DimensionMismatch("camera pixel vectors do not match the screen size")
        c.size = size if size is not None else (len(c.alphas), 1)
        if c.size[0] * c.size[1] != len(c.alphas):
            raise DimensionMismatch("camera pixel vectors do not match the screen size")
        return c

    @property
    def npix(self):
        return len(self.alphas)


def tiles(camera, ntiles):
    """tiles(camera, ntiles) (cache.jl:234-246): `ntiles` cameras of consecutive pixels (screen order)
    with the index range each covers, so that a screen too large for stored samples is regenerated
    tile by tile with a smaller cache and the results written back into the full screen."""
    n = camera.npix
    bounds = np.round(np.linspace(0, n, int(ntiles) + 1)).astype(np.int64)
    return [(Camera.points(camera.alphas[b0:b1], camera.betas[b0:b1]), range(int(b0), int(b1)))
            for b0, b1 in zip(bounds[:-1], bounds[1:]) if b1 > b0]


class Binning:
    """Binning(pixel, size) (camera.jl:62-80): pixel integration of a camera whose points sub-sample
    pixels. pixel[m] is the pixel of point m (screen order), a 0-based linear index i + n_alpha j into
    a screen of shape size = (n_alpha, n_beta) (the reference's 1-based index minus one); a pixel's
    value is the mean over its points (`bin`). Every pixel needs at least one point."""

    def __init__(self, pixel, size):
        self.pixel = np.asarray(pixel, dtype=np.int64).ravel()
        self.size = (int(size[0]), int(size[1]))
        npix = self.size[0] * self.size[1]
        bad = (self.pixel < 0) | (self.pixel >= npix)
        if bad.any():
            raise ValueError(f"pixel index {int(self.pixel[bad][0])} outside 0:{npix - 1}")
        self.count = np.bincount(self.pixel, minlength=npix)
        if (self.count == 0).any():
            raise ValueError("every pixel needs at least one point")
        self._h = None

    @property
    def npixels(self):
        return self.size[0] * self.size[1]

    @property
    def npoints(self):
        return len(self.pixel)

    def handle(self, device=0):
        """The binning on the device (ks_binning_create), made on first use."""
        if self._h is None:
            self._lib = _load()
            h = _P()
            _check(self._lib.ks_binning_create(context(device), _i64(self.npoints), _ptr(self.pixel), _i(self.size[0]), _i(self.size[1]),
                                               ctypes.byref(h)))
            self._h = h
        return self._h

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                self._lib.ks_binning_destroy(self._h)
                self._h = None
        except Exception:
            pass


def bin(binning, values):
    """bin(binning, values) (camera.jl:82-105): the mean of the values of each pixel's points. values
    has one entry per point along its leading axis (screen order; trailing axes, e.g. Stokes, are kept),
    or is a cube (npoints, 1, nt, nnu[, 4]) as polarized_cube returns for a point camera; returns
    (n_alpha, n_beta, ...)."""
    x = np.asarray(values, dtype=np.float64)
    if x.ndim >= 2 and x.shape[:2] == (binning.npoints, 1):
        x = x.reshape((binning.npoints,) + x.shape[2:])
    if x.shape[0] != binning.npoints:
        raise DimensionMismatch(f"expected {binning.npoints} values, got {x.shape[0]}")
    acc = np.zeros((binning.npixels,) + x.shape[1:])
    np.add.at(acc, binning.pixel, x)                  # in point order, as the reference's loop
    acc /= binning.count.reshape((-1,) + (1,) * (x.ndim - 1))
    na, nb = binning.size
    return acc.reshape((nb, na) + x.shape[1:]).swapaxes(0, 1)


def binned_grid(arange, brange, size, subsamples=1):
    """binned_grid((alo, ahi), (blo, bhi), (na, nb); subsamples) (camera.jl:107-128): na x nb square pixels
    covering the given edges, each integrated over subsamples^2 points at the centres of its sub-cells;
    points listed pixel by pixel in screen order. Returns (camera, binning)."""
    na, nb = (int(size), int(size)) if np.ndim(size) == 0 else (int(size[0]), int(size[1]))
    K = int(subsamples)
    if K < 1:
        raise ValueError("subsamples must be at least 1")
    da, db = (arange[1] - arange[0]) / na, (brange[1] - brange[0]) / nb
    s = (np.arange(1, K + 1) - 0.5) / K
    j, i, sj, si = np.meshgrid(np.arange(nb), np.arange(na), s, s, indexing="ij")     # j slowest, si fastest
    al = arange[0] + (i + si) * da
    be = brange[0] + (j + sj) * db
    pixel = i + j * na
    return Camera.points(al.ravel(), be.ravel()), Binning(pixel.ravel(), (na, nb))


def binned_polar(rho_edges, npsi, subsamples=(1, 1), psi0=0.0):
    """binned_polar(ρedges, nψ; subsamples, ψ0) (camera.jl:130-149): an annulus of polar pixels, radial
    cells between consecutive edges and npsi azimuthal cells from psi0, pixel (iρ, iψ) at index
    iρ + nρ iψ, each integrated over subsamples[0] x subsamples[1] (ρ, ψ) sub-cells (α = ρ cos ψ,
    β = ρ sin ψ). Returns (camera, binning)."""
    edges = np.asarray(rho_edges, dtype=np.float64)
    nr = len(edges) - 1
    if nr < 1:
        raise ValueError("need at least two radial edges")
    Kr, Kp = int(subsamples[0]), int(subsamples[1])
    dpsi = 2 * np.pi / int(npsi)
    ip, ir, sp, sr = np.meshgrid(np.arange(int(npsi)), np.arange(nr), np.arange(1, Kp + 1), np.arange(1, Kr + 1), indexing="ij")
    rho = edges[ir] + (sr - 0.5) / Kr * (edges[ir + 1] - edges[ir])
    psi = psi0 + (ip + (sp - 0.5) / Kp) * dpsi
    return Camera.points((rho * np.cos(psi)).ravel(), (rho * np.sin(psi)).ravel()), Binning((ir + ip * nr).ravel(), (nr, int(npsi)))


def concatenate(*parts):
    """concatenate((camera, binning), ...) (camera.jl:151-164): the points and the pixels in order (the
    pixels of the second screen follow those of the first), binned screen shape (npixels, 1)."""
    al = np.concatenate([c.alphas for c, _ in parts])
    be = np.concatenate([c.betas for c, _ in parts])
    pixel, offset = [], 0
    for _, b in parts:
        pixel.append(b.pixel + offset)
        offset += b.npixels
    return Camera.points(al, be), Binning(np.concatenate(pixel), (offset, 1))


def concatenate_screens(*cameras):
    """One point camera from several screens (concatenate, camera.jl:151-164, without a binning): the points of
    each camera in order, each camera kept as a screen part (its size) so that the device groups the work of
    one screen patch by patch (set_parts). Returns the camera; camera.parts lists the part sizes."""
    cam = Camera.points(np.concatenate([c.alphas for c in cameras]), np.concatenate([c.betas for c in cameras]))
    cam.parts = [tuple(int(x) for x in c.size) for c in cameras]
    return cam


def screen_tiles(camera, splits):
    """The screen of a grid camera cut into rectangular tiles: splits = [(a0, a1, b0, b1), ...] pixel index
    ranges (half-open) that cover the screen once. Returns (cameras, index lists): each tile as a grid camera
    over its own pixels (Krang's grid positions of the full screen) and the screen indices of its pixels in
    the tile's screen order (the reference's tiles, cache.jl:234-246, cut runs of consecutive pixels instead)."""
    na, nb = camera.size
    A = camera.alphas.reshape(na, nb, order="F")
    B = camera.betas.reshape(na, nb, order="F")
    cams, idx, seen = [], [], np.zeros((na, nb), dtype=np.int64)
    for a0, a1, b0, b1 in splits:
        c = Camera.points(A[a0:a1, b0:b1].ravel(order="F"), B[a0:a1, b0:b1].ravel(order="F"), size=(a1 - a0, b1 - b0))
        cams.append(c)
        ia, ib = np.meshgrid(np.arange(a0, a1), np.arange(b0, b1), indexing="ij")
        idx.append((ia + na * ib).ravel(order="F"))
        seen[a0:a1, b0:b1] += 1
    if not (seen == 1).all():
        raise ValueError("the tiles must cover the screen exactly once")
    return cams, idx


class Direct:
    code, M = KS_MARCHER_DIRECT, 64


class Recurrence:
    code = KS_MARCHER_RECURRENCE

    def __init__(self, M=64):
        self.M = int(M)


class Fused:
    code = KS_MARCHER_FUSED

    def __init__(self, M=64):
        self.M = int(M)


class GeodesicCache:
    """GeodesicCache(camera, N; store_samples) on the GPU (cache.jl:85-125)."""

    def __init__(self, camera, N, store_samples=True, device=0):
        if N < 1:
            raise ValueError("need at least one sample per ray")
        self._lib = _load()
        self._ctx = context(device)
        self.camera = camera
        self.N = int(N)
        self.store_samples = bool(store_samples)
        p = _P()
        _check(self._lib.ks_cache_create(self._ctx, _i64(camera.npix), _i(self.N), _i(int(self.store_samples)), ctypes.byref(p)))
        self._h = p
        self.spin = math.nan
        self.theta_o = math.nan
        self.generation = 0
        self.marcher = Direct()
        self.perm_host = np.arange(1, camera.npix + 1)
        self.ranges = (0, 0, 0)
        self.precision = "fp64"
        _set_parts(self, camera)

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                self._lib.ks_cache_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def npix(self):
        return self.camera.npix

    @property
    def screen_size(self):
        return self.camera.size

    def info(self):
        npix, N, gen = _i64(), _i(), _i()
        counts = (_i64 * 3)()
        _check(self._lib.ks_cache_info(self._h, ctypes.byref(npix), ctypes.byref(N), ctypes.byref(gen), counts))
        return npix.value, N.value, gen.value, tuple(counts)


def regenerate(cache, a, theta_o, camera=None, marcher=None):
    """regenerate!(cache, a, θo[, camera]; marcher) (cache.jl:153-180, 248-255)."""
    marcher = marcher if marcher is not None else Direct()
    if camera is not None:
        if camera.npix != cache.npix:
            raise DimensionMismatch(f"camera has {camera.npix} pixels, cache was built for {cache.npix}")
        cache.camera = camera
        _set_parts(cache, camera)
    cam = cache.camera
    al, be = _f64(cam.alphas), _f64(cam.betas)
    lib = cache._lib
    _check(lib.ks_regenerate(cache._h, _ptr(al), _ptr(be), _i(cam.size[0]), _i(cam.size[1]), _d(a), _d(theta_o),
                             _i(marcher.code), _i(getattr(marcher, "M", 64))))
    perm = np.zeros(cache.npix, dtype=np.int64)
    counts = np.zeros(3, dtype=np.int64)
    _check(lib.ks_get_perm(cache._h, _ptr(perm), _ptr(counts)))
    cache.perm_host = perm
    cache.ranges = tuple(int(c) for c in counts)
    cache.spin, cache.theta_o = float(a), float(theta_o)
    cache.generation += 1
    cache.marcher = marcher
    return cache


def _set_parts(cache, camera):
    parts = getattr(camera, "parts", None) or []
    na = np.array([p[0] for p in parts], dtype=np.int32)
    nb = np.array([p[1] for p in parts], dtype=np.int32)
    _check(cache._lib.ks_cache_set_parts(cache._h, _i(len(parts)), _ptr(na) if len(parts) else None, _ptr(nb) if len(parts) else None))


def tuning(cache, chunk=None, queue=None, patch=None, adjoint_queue=None):
    """The launch choices of the cache's transfer over stored samples and of its element adjoints: samples per work
    item (chunk), the transfer's persistent work-queue form (queue), the element adjoints' persistent form
    (adjoint_queue), the wavefront's screen patch (patch = (n_alpha, n_beta) points); arguments left at None keep
    their value. Returns the current choices as a dict."""
    lib = cache._lib
    c, q, pa, pb = _i(), _i(), _i(), _i()
    _check(lib.ks_cache_get_tuning(cache._h, ctypes.byref(c), ctypes.byref(q), ctypes.byref(pa), ctypes.byref(pb)))
    if chunk is not None or queue is not None or patch is not None or adjoint_queue is not None:
        nc = int(chunk) if chunk is not None else c.value
        tq = bool(queue) if queue is not None else bool(q.value & KS_QUEUE_TRANSFER)
        aq = bool(adjoint_queue) if adjoint_queue is not None else bool(q.value & KS_QUEUE_ADJOINT)
        npa, npb = (int(patch[0]), int(patch[1])) if patch is not None else (pa.value, pb.value)
        _check(lib.ks_cache_tuning(cache._h, _i(nc), _i((KS_QUEUE_TRANSFER if tq else 0) | (KS_QUEUE_ADJOINT if aq else 0)), _i(npa), _i(npb)))
        _check(lib.ks_cache_get_tuning(cache._h, ctypes.byref(c), ctypes.byref(q), ctypes.byref(pa), ctypes.byref(pb)))
    return {"chunk": c.value, "queue": bool(q.value & KS_QUEUE_TRANSFER), "patch": (pa.value, pb.value),
            "adjoint_queue": bool(q.value & KS_QUEUE_ADJOINT)}


def autotune(cache, model, times, nus, L, records=True, sample_every=16, adjoint=False, transfer=True):
    """The autotuner: times every launch choice of the transfer (chunk 16/32/64/128 x patch 8x8/16x4/4x16 x queue
    off/on) on a hashed sample of 1 / sample_every of the transfer's tasks for (times, nus) (records = the dual
    sweep's records written, as in gradient_forward) and keeps the fastest on the cache (transfer=False keeps the
    cache's transfer choices); with adjoint, also times one reverse pass in each launch form of the element adjoints
    and keeps the faster. Returns (choice, table), table = [(chunk, (patch_a, patch_b), queue bits, ms), ...] (the
    element adjoints' rows with chunk 0)."""
    m = _as_model(model, "thermal")
    tv, nv = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    cap = 64
    table = np.zeros(5 * cap)
    n = _i()
    flags = (KS_TUNE_RECORDS if records else 0) | (KS_TUNE_ADJOINT if adjoint else 0) | (0 if transfer else KS_TUNE_KEEP_TRANSFER)
    _check(cache._lib.ks_autotune(cache._h, m._h, _ptr(tv), _i(len(tv)), _ptr(nv), _i(len(nv)), _d(L), _i(flags),
                                  _i(int(sample_every)), _ptr(table), _i(cap), ctypes.byref(n)))
    rows = [(int(r[0]), (int(r[1]), int(r[2])), int(r[3]), float(r[4])) for r in table[:5 * min(n.value, cap)].reshape(-1, 5)]
    return tuning(cache), rows


def queue_costs(cache):
    """The per-task wall-clock ticks (100 MHz) of the cache's last work-queue launch, indexed by task."""
    n = _i64()
    _check(cache._lib.ks_queue_costs(cache._h, ctypes.byref(n), None, _i(0)))
    out = np.zeros(n.value, dtype=np.uint32)
    _check(cache._lib.ks_queue_costs(cache._h, ctypes.byref(n), _ptr(out), _i(len(out))))
    return out


class FitGraph:
    """A fit iteration captured as one hipGraph (ks_fit_graph_*): the upload of the splat constants, the transfer with
    the records over every (frame, frequency) slice, the images, every slice's chi2 and seed, the reverse pass.
    FitGraph(cache, times, nus, L, image=(data, sigma[, mask])) with data (n_alpha, n_beta, nt, nnu, 4) in screen
    order (or (npix, nt, nnu, 4)) and sigma 4 numbers or an array of data's shape, or vis=(psize, u, v, data, sigma[,
    taper]) with data (nt, nnu, nuv, 4) complex and sigma (nt, nnu, nuv, 4). step(model, params=None) -> (chi2,
    gradient): the model's constants remade from params (its shape; None: the model's current constants), one replay
    (captured again for another model, e.g. after a hygiene pass changed the splat count)."""

    def __init__(self, cache, times, nus, L, image=None, vis=None):
        if (image is None) == (vis is None):
            raise ValueError("FitGraph takes one loss: image=(data, sigma[, mask]) or vis=(psize, u, v, data, sigma[, taper])")
        self._lib, self.cache = cache._lib, cache
        tv, nv = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
        nt, nnu, n = len(tv), len(nv), cache.npix
        self.frames = nt * nnu
        h = _P()
        if image is not None:
            data, sigma = image[0], image[1]
            mask = image[2] if len(image) > 2 else None
            d = np.asarray(data, dtype=np.float64).reshape(n, nt, nnu, 4, order="F")
            dd = np.ascontiguousarray(np.stack([d[:, it, inu, :] for inu in range(nnu) for it in range(nt)]))
            sg = np.asarray(sigma, dtype=np.float64)
            per_pixel = sg.size > 4
            if per_pixel:
                sg = sg.reshape(n, nt, nnu, 4, order="F")
                sd = np.ascontiguousarray(np.stack([sg[:, it, inu, :] for inu in range(nnu) for it in range(nt)]))
            else:
                sd = np.ascontiguousarray(np.tile(sg.ravel(), self.frames))
            md = None
            if mask is not None:
                mk = np.asarray(mask, dtype=bool).reshape(n, nt, nnu, order="F")
                md = np.ascontiguousarray(np.stack([mk[:, it, inu] for inu in range(nnu) for it in range(nt)]).astype(np.uint8))
            self._keep = (tv, nv, dd, sd, md)
            _check(self._lib.ks_fit_graph_create(cache._h, _i(nt), _ptr(tv), _i(nnu), _ptr(nv), _d(L), _i(0), _ptr(dd), _ptr(sd),
                                                 _i(int(per_pixel)), _ptr(md), _d(0.0), None, None, None, _i(0), ctypes.byref(h)))
        else:
            psize, u, v, data, sigma = vis[:5]
            taper = vis[5] if len(vis) > 5 else None
            uu, vv = _f64(u), _f64(v)
            nuv = len(uu)
            D = np.asarray(data).reshape(nt, nnu, nuv, 4)
            S = np.asarray(sigma, dtype=np.float64).reshape(nt, nnu, nuv, 4)
            dd = np.ascontiguousarray(np.stack([np.stack([D[it, inu].real, D[it, inu].imag], axis=-1) for inu in range(nnu) for it in range(nt)]))
            sd = np.ascontiguousarray(np.stack([S[it, inu] for inu in range(nnu) for it in range(nt)]))
            tp = _f64(taper) if taper is not None else None
            self._keep = (tv, nv, uu, vv, dd, sd, tp)
            _check(self._lib.ks_fit_graph_create(cache._h, _i(nt), _ptr(tv), _i(nnu), _ptr(nv), _d(L), _i(1), _ptr(dd), _ptr(sd), _i(0),
                                                 None, _d(psize), _ptr(uu), _ptr(vv), _ptr(tp), _i(nuv), ctypes.byref(h)))
        self._h = h

    def step(self, model, params=None):
        if model.kind not in ("thermal", "powerlaw", "kappa", "knot"):
            raise ValueError("a fit graph takes one thermal, power-law, kappa or knot population")
        P = None
        if params is not None:
            P = np.asfortranarray(params, dtype=np.float64)
            if P.shape != model.grad_shape:
                raise DimensionMismatch(f"params of shape {P.shape}, the model holds {model.grad_shape}")
            model.params = P
        chi = _d()
        g = np.zeros(model.grad_shape, order="F")
        _check(self._lib.ks_fit_graph_step(self._h, model._h, _ptr(P), ctypes.byref(chi), _ptr(g)))
        return chi.value, g

    def info(self):
        c, r = _i64(), _i64()
        _check(self._lib.ks_fit_graph_info(self._h, ctypes.byref(c), ctypes.byref(r)))
        return {"captures": c.value, "replays": r.value}

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                self._lib.ks_fit_graph_destroy(self._h)
                self._h = None
        except Exception:
            pass


def unsort(cache, x):
    """unsort(cache, x): sorted-order vector/matrix (first axis npix) back to screen order (cache.jl:271-286)."""
    x = np.asarray(x)
    if x.shape[0] != cache.npix:
        raise DimensionMismatch("first dimension must be the pixel count")
    out = np.empty_like(x)
    out[cache.perm_host - 1] = x
    return out


def to_screen(cache, x):
    """to_screen(cache, x) (cache.jl:293-296): unsort and reshape to (n_alpha, n_beta, ...)."""
    y = unsort(cache, x)
    return y.reshape(cache.screen_size + y.shape[1:], order="F")


def samples(cache):
    """host(cache.samples) in sorted order: dict of (npix, N) arrays (direct_march.jl:71-115)."""
    n, N = cache.npix, cache.N
    t, r, th, ph = (np.zeros((N, n)) for _ in range(4))
    fl = np.zeros((N, n), dtype=np.uint8)
    _check(cache._lib.ks_copy_samples(cache._h, _ptr(t), _ptr(r), _ptr(th), _ptr(ph), _ptr(fl)))
    return {"t": t.T.copy(), "r": r.T.copy(), "theta": th.T.copy(), "phi": ph.T.copy(), "flags": fl.T.copy()}


PIXEL_FIELDS = ("alpha", "beta", "lam", "eta", "case", "k", "c", "fo", "tau_total", "tau_h", "r1", "r3", "r4", "n", "n0", "n1",
                "d0", "d1", "a2", "b1", "g0", "psi0", "Pp", "Pm", "Qp", "Qm", "It0", "Iph0", "up", "um", "f", "mth", "off", "sgn",
                "npol", "Xth0", "Gt0", "Gphi0", "vortical", "mu", "scale", "K2", "nreal")


def consts(cache):
    """host(cache.consts) (Geodesics.jl:66-73 over PixelConstants, pixel_constants.jl:29-62): the per-pixel
    constants in the cache's sorted order as a dict of (npix,) arrays named as in PIXEL_FIELDS."""
    nf = cache._lib.ks_num_const_fields()
    if nf != len(PIXEL_FIELDS):
        raise KernelError(f"the library holds {nf} constants per pixel, this module names {len(PIXEL_FIELDS)}")
    soa = np.zeros((nf, cache.npix))
    _check(cache._lib.ks_copy_consts(cache._h, _ptr(soa)))
    return {name: soa[i].copy() for i, name in enumerate(PIXEL_FIELDS)}


def residuals(cache):
    rt, rp = np.zeros(cache.npix), np.zeros(cache.npix)
    _check(cache._lib.ks_get_residuals(cache._h, _ptr(rt), _ptr(rp)))
    return rt, rp


class Model:
    """A splat parameter matrix on the device: ThinRenderer / PolarizedSplats / PowerLawSplats /
    KappaSplats / KnotSplats (Splats.jl:105, polarized.jl:41, populations.jl:34-60, trajectories.jl:80-88)."""

    def __init__(self, params, kind="thermal", gamma_max=1e5, knots=None, knot_times=None, device=0):
        if kind not in MODEL_KIND:
            raise ValueError(f"unknown model kind {kind}")
        P = np.asfortranarray(params, dtype=np.float64)
        if P.ndim != 2 or P.shape[0] != _NROWS[kind]:
            raise DimensionMismatch(f"{kind} splats take {_NROWS[kind]} parameter rows, got {P.shape}")
        self.kind, self.params = kind, P
        self._lib = _load()
        kn = np.asfortranarray(knots, dtype=np.float64) if knots is not None else None
        kt = _f64(knot_times) if knot_times is not None else None
        h = _P()
        _check(self._lib.ks_model_create(context(device), _i(MODEL_KIND[kind]), _i(P.shape[0]), _i(P.shape[1]), _ptr(P),
                                         _d(gamma_max), _ptr(kn), _ptr(kt), _i(0 if kt is None else len(kt)), ctypes.byref(h)))
        self._h = h
        self.ns = P.shape[1]
        self.grad_shape = P.shape

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                self._lib.ks_model_destroy(self._h)
                self._h = None
        except Exception:
            pass


class CompositeModel(Model):
    """CompositeModel(a, b, ...) (transport.jl:298-316): the elements of a, then of b, ...; the
    coefficients of overlapping elements add. Parts are Model objects of kind thermal, powerlaw, kappa
    or knot (each knot part with its own knots and knot times), or other composites; the splat constants
    and knots are copied to the composite on the device."""

    def __init__(self, *parts, device=0):
        if not parts:
            raise ValueError("CompositeModel needs at least one part")
        for p in parts:
            if not isinstance(p, Model) or p.kind not in ("thermal", "powerlaw", "kappa", "knot", "composite"):
                raise ValueError("CompositeModel parts are thermal, powerlaw, kappa, knot or composite Model objects")
        self.kind, self.params, self.parts = "composite", None, parts
        self._lib = _load()
        arr = (_P * len(parts))(*[p._h.value for p in parts])
        h = _P()
        _check(self._lib.ks_model_composite(context(device), _i(len(parts)), arr, ctypes.byref(h)))
        self._h = h
        self.ns = sum(p.ns for p in parts)
        # the gradient of a composite: its parts' nrows x ns blocks (column-major) one after another
        self.grad_shape = (sum(int(np.prod(p.grad_shape)) for p in parts),)

    def split_gradient(self, g):
        """The parts' gradient blocks (each shaped like its part's parameters) of a composite gradient vector."""
        out, off = [], 0
        for p in self.parts:
            n = int(np.prod(p.grad_shape))
            blk = np.asarray(g)[off:off + n]
            out.extend(p.split_gradient(blk) if isinstance(p, CompositeModel) else [blk.reshape(p.grad_shape, order="F")])
            off += n
        return out


def _as_model(params, kind):
    return params if isinstance(params, Model) else Model(params, kind=kind)


def thin_image(cache, params, t_obs, screen=True):
    """thin_image(cache, params, t_obs) (Splats.jl:137-147): screen-shaped (or sorted) Stokes I."""
    m = _as_model(params, "thin")
    out = np.zeros(cache.npix)
    _check(cache._lib.ks_thin_image(cache._h, m._h, _d(t_obs), _ptr(out), _i(int(screen))))
    return out.reshape(cache.screen_size, order="F") if screen else out


class RayLists:
    """RayLists(ids, count) (polarized.jl:53-75): for every ray the splats whose support its samples enter.
    ids is (capacity, npix) or (capacity, npix, nframes) int32, 0-based splat indices in ascending order
    with -1 past each ray's count; count is (npix,) or (npix, nframes); rays in the camera's pixel order."""

    def __init__(self, ids, count):
        self.ids, self.count = np.asarray(ids, dtype=np.int32), np.asarray(count, dtype=np.int32)
        if self.ids.shape[1:] != self.count.shape:
            raise DimensionMismatch(f"ids {self.ids.shape} and count {self.count.shape} disagree")

    @property
    def capacity(self):
        return self.ids.shape[0]

    @property
    def nframes(self):
        return self.ids.shape[2] if self.ids.ndim == 3 else 1

    def ray(self, j, f=0):
        """The splats of ray j in frame f."""
        ids = self.ids if self.ids.ndim == 3 else self.ids[:, :, None]
        cnt = self.count if self.count.ndim == 2 else self.count[:, None]
        return ids[: cnt[j, f], j, f]


def capacity(lists):
    return lists.capacity


def nframes(lists):
    return lists.nframes


def framed(lists):
    """The lists as (capacity, npix, nframes) and (npix, nframes) (polarized.jl:68-70)."""
    if lists is None or lists.ids.ndim == 3:
        return lists
    return RayLists(lists.ids[:, :, None], lists.count[:, None])


def flattened(lists):
    """The lists flattened over frames: (capacity, npix nframes) and a vector, frame-major (polarized.jl:71-72)."""
    return RayLists(lists.ids.reshape(lists.capacity, -1, order="F"), lists.count.reshape(-1, order="F"))


def frame_lists(lists, f):
    """The lists of frame f of framed lists (polarized.jl:73-75)."""
    if lists is None:
        return None
    return RayLists(lists.ids[:, :, f], lists.count[:, f])


def ray_lists(cache, params, times, nmax=-1, slab=0.0, kind="thermal"):
    """ray_lists(cache, params, t_obs | times; nmax, slab) -> RayLists (dual_gradient.jl:326-359): for every
    ray the splats whose support sphere one of its step-taking samples enters at the observation time,
    samples past the half-orbit cutoff excluded, in ascending order; built on the device (a hit byte per
    splat and ray, the counts, then the lists filled into capacity = the largest count slots). A scalar
    time gives (capacity, npix) ids and (npix,) counts, a vector of times (capacity, npix, nframes) and
    (npix, nframes). The transport over the lists is the same sum as over all splats; this build's image
    and gradient launches cull per tile of samples on the device and take no lists."""
    m = _as_model(params, kind)
    scalar = np.ndim(times) == 0
    t = _f64(np.atleast_1d(times))
    npix, nt = cache.npix, len(t)
    count = np.zeros((nt, npix), dtype=np.int32)
    cap = _i(0)
    _check(cache._lib.ks_ray_lists(cache._h, m._h, _ptr(t), _i(nt), _i(nmax), _d(slab), _i(1), _ptr(count), ctypes.byref(cap), None))
    ids = np.zeros((nt, npix, cap.value), dtype=np.int32)
    _check(cache._lib.ks_ray_lists(cache._h, m._h, _ptr(t), _i(nt), _i(nmax), _d(slab), _i(1), _ptr(count), ctypes.byref(cap), _ptr(ids)))
    ids, count = np.transpose(ids, (2, 1, 0)), count.T
    return RayLists(ids[:, :, 0], count[:, 0]) if scalar else RayLists(ids, count)


def render_device(cache, model, times, nus, L, nmax=-1, slab=0.0):
    """The fused polarized launch with the result left on the device (ks_render_device)."""
    t, n = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    _check(cache._lib.ks_render_device(cache._h, model._h, _ptr(t), _i(len(t)), _ptr(n), _i(len(n)), _d(L), _i(nmax), _d(slab)))


def spacetime_tangent(cache, params, times, nus, L, nmax=-1, slab=0.0, kind="thermal", screen=True):
    """The observed Stokes cube of the cache's camera at its spin and inclination and its forward-mode partials
    with respect to x = (a, theta_o, ln L) (the dual cache of spacetime_valgrad / spacetime_jacobian,
    joint.jl:74-154): three-partial duals through the pixel constants, the closed-form samples and the
    transport. Returns (cube, dcube): (n_alpha, n_beta, nt, nnu, 4) and (n_alpha, n_beta, nt, nnu, 4, 3) with
    screen=True, (nnu, nt, npix, 4[, 3]) in screen pixel order otherwise."""
    m = _as_model(params, kind)
    t, n = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    out = np.zeros((len(n), len(t), cache.npix, 4))
    dout = np.zeros((len(n), len(t), cache.npix, 4, 3))
    _check(cache._lib.ks_spacetime_tangent(cache._h, m._h, _ptr(t), _i(len(t)), _ptr(n), _i(len(n)), _d(L), _i(nmax), _d(slab),
                                           _ptr(out), _ptr(dout)))
    if not screen:
        return out, dout
    na, nb = cache.screen_size
    return (np.transpose(out.reshape(len(n), len(t), nb, na, 4), (3, 2, 1, 0, 4)),
            np.transpose(dout.reshape(len(n), len(t), nb, na, 4, 3), (3, 2, 1, 0, 4, 5)))


def spacetime_tangent_differences(camera, N, a, theta_o, params, times, nus, L, nmax=-1, slab=0.0, kind="thermal", h=2e-6,
                                  screen=True, marcher=None):
    """The observed Stokes cube of `camera` at (a, theta_o) with N stored samples per ray (marcher: Recurrence(64),
    the spacetime block's stored march, joint.jl:78, by default; Direct() for the closed-form samples; Fused(64), the
unstored form, marches inside each render) and its
    partials with respect to x = (a, theta_o, ln L) (the values spacetime_valgrad / spacetime_jacobian take from their
    dual cache, joint.jl:74-154), every image rendered on the device in FP64: central differences in each variable,
    Richardson-extrapolated from the steps h and h/2 (truncation error O(h^4); the reference checks its duals
    against such differences, test_spacetime.jl:94-113). Steps much longer than 1e-6 can straddle a sample that
    crosses the horizon guard or a splat's weight cutoff, a jump that the partials exclude. The spin and
    inclination partials regenerate the pixel constants and samples at a +- h and theta_o +- h (8 marches); the
    path-length partial re-renders the cube at L e^(+-h) (13 cubes in all). Returns (cube, dcube) shaped as
    spacetime_tangent's."""
    if not (L > 0.0):
        raise ValueError("the length unit L must be positive")
    if not (h > 0.0):
        raise ValueError("the step h must be positive")
    m = _as_model(params, kind)
    march = Recurrence(64) if marcher is None else marcher
    cache = GeodesicCache(camera, N, store_samples=not isinstance(march, Fused))

    def cube(a_, th_, L_):
        regenerate(cache, a_, th_, marcher=march)
        return polarized_cube(cache, m, times, nus, L_, nmax=nmax, slab=slab, screen=screen)

    def central(k, hh):
        e = np.eye(3)[k] * hh
        up = cube(a + e[0], theta_o + e[1], L * math.exp(e[2]))
        dn = cube(a - e[0], theta_o - e[1], L * math.exp(-e[2]))
        return (up - dn) / (2.0 * hh)

    dcube = np.stack([(4.0 * central(k, 0.5 * h) - central(k, h)) / 3.0 for k in range(3)], axis=-1)
    return cube(a, theta_o, L), dcube


def polarized_cube(cache, params, times, nus, L, nmax=-1, slab=0.0, kind="thermal", screen=True):
    """polarized_cube(cache, params, times, νs, L; nmax, slab) (polarized.jl:187-198), all frames
    and frequencies in one launch: (n_alpha, n_beta, nt, nnu, 4) observed Stokes [cgs]."""
    m = _as_model(params, kind)
    t, n = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    out = np.zeros((len(n), len(t), cache.npix, 4))
    _check(cache._lib.ks_render(cache._h, m._h, _ptr(t), _i(len(t)), _ptr(n), _i(len(n)), _d(L), _i(nmax), _d(slab),
                                _ptr(out), _i(int(screen))))
    if not screen:
        return out
    na, nb = cache.screen_size
    img = out.reshape(len(n), len(t), nb, na, 4)            # pixel index = i + na j
    return np.transpose(img, (3, 2, 1, 0, 4))


def polarized_image(cache, params, t_obs, nu, L, nmax=-1, slab=0.0, kind="thermal", screen=True):
    """polarized_image(cache, params, t_obs, ν, L; nmax, slab) (polarized.jl:174-179):
    (n_alpha, n_beta, 4) observed Stokes (or (npix, 4) in sorted order with screen=False)."""
    c = polarized_cube(cache, params, [t_obs], [nu], L, nmax=nmax, slab=slab, kind=kind, screen=screen)
    return c[:, :, 0, 0, :] if screen else c[0, 0]


STAGES = {"transfer": 0, "adjoint_steps": 1, "adjoint_elements": 2, "march": 3, "vis": 4}


def stage_ms(stage, device=0):
    """Duration [ms] of the last launch of a pipeline stage ("transfer", "adjoint_steps",
    "adjoint_elements", "march", "vis": the visibility transform's matrix-core product) on the device's
    context, from HIP events on its stream."""
    ms = ctypes.c_float()
    _check(_load().ks_stage_ms(context(device), _i(STAGES[stage]), ctypes.byref(ms)))
    return ms.value


PRECISIONS = {"fp64": 0, "mixed": 1}


def precision(cache, policy):
    """precision(cache, T2) (cache.jl:212-232): "fp64" (default) or "mixed" -- the transfer and the
    dual sweep over the stored samples in FP32 (FP64 geodesics and per-sample geometry; the step in
    FP64 where the Faraday depth exceeds 10; compensated Stokes sums). Set in place; returns the cache."""
    if policy not in PRECISIONS:
        raise ValueError(f"precision policy must be one of {sorted(PRECISIONS)}, got {policy!r}")
    _check(cache._lib.ks_cache_precision(cache._h, _i(PRECISIONS[policy])))
    cache.precision = policy
    return cache


def march_precision(cache, policy):
    """The scalar of the Recurrence(M) march into stored samples for the next regenerate / march: "fp64" (default) or
    "mixed" (the integrand rates and Simpson increments of t~ and phi in FP32 on the well-conditioned segments; FP64
    Jacobi states, positions, anchors, root cases, pixel constants, Mino times, running sums and stored samples).
    Set in place; returns the cache."""
    if policy not in PRECISIONS:
        raise ValueError(f"precision policy must be one of {sorted(PRECISIONS)}, got {policy!r}")
    _check(cache._lib.ks_cache_march_precision(cache._h, _i(PRECISIONS[policy])))
    cache.march_precision = policy
    return cache


def march(cache):
    """march!(marcher, cache) (cache.jl:182-190): re-run the cache's storing marcher into its samples."""
    _check(cache._lib.ks_march(cache._h))
    return cache


def polarized_gradient(cache, params, times, nus, L, dstokes, dparams=None, nmax=-1, slab=0.0, screen=True):
    """polarized_gradient!(dparams, dstokes, cache, params, t_obs, ν, L; method = :dual, nmax, slab)
    (polarized.jl:473-480, dual_gradient.jl:180-249): the gradient of sum(dstokes * I_obs) over every
    pixel, time and frequency with respect to the splat parameters (thermal, power-law, kappa, knot or composite
    models), added into dparams (nrows x ns; for a composite the parts' blocks one after another as a vector,
    split by CompositeModel.split_gradient; allocated as zeros when None). dstokes has the shape polarized_cube returns
    ((n_alpha, n_beta, nt, nnu, 4) with screen=True, (nnu, nt, npix, 4) in sorted order otherwise);
    returns (dparams, cube) with the cube of the same forward pass. Needs stored samples."""
    m = _as_model(params, "thermal")
    if m.kind == "thin":
        raise ValueError("the dual sweep takes polarized splats (thermal, power-law, kappa, knot or composite)")
    t, n = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    na, nb = cache.screen_size
    ds = np.asarray(dstokes, dtype=np.float64)
    if screen:
        if ds.shape != (na, nb, len(t), len(n), 4):
            raise DimensionMismatch(f"dstokes must be {(na, nb, len(t), len(n), 4)}, got {ds.shape}")
        # (n_alpha, n_beta, nt, nnu, 4) -> (nnu, nt, npix, 4) with pixel = i + na j
        ds = np.ascontiguousarray(np.transpose(ds, (3, 2, 1, 0, 4)).reshape(len(n), len(t), cache.npix, 4))
    else:
        if ds.shape != (len(n), len(t), cache.npix, 4):
            raise DimensionMismatch(f"dstokes must be {(len(n), len(t), cache.npix, 4)}, got {ds.shape}")
        ds = np.ascontiguousarray(ds)
    if dparams is None:
        dparams = np.zeros(m.grad_shape, order="F")
    elif dparams.shape != m.grad_shape or not dparams.flags.f_contiguous or dparams.dtype != np.float64:
        raise DimensionMismatch("dparams must be a Fortran-ordered float64 array shaped like the parameters (a vector for composites)")
    out = np.zeros((len(n), len(t), cache.npix, 4))
    _check(cache._lib.ks_polarized_gradient(cache._h, m._h, _ptr(t), _i(len(t)), _ptr(n), _i(len(n)), _d(L), _i(nmax), _d(slab),
                                            _ptr(ds), _i(int(screen)), _ptr(dparams), _ptr(out)))
    if not screen:
        return dparams, out
    img = out.reshape(len(n), len(t), nb, na, 4)
    return dparams, np.transpose(img, (3, 2, 1, 0, 4))


def visibilities(cache, frames, psize, u, v, taper=None, noalloc=False):
    """Visibilities of the rendered frames (visibilities.jl:21-40, device_visibilities.jl:61-73):
    complex (frames, nuv, 4) in Jy for pixel size psize = Δα L / D [rad]. noalloc takes the entry point
    that allocates no device memory (its workspace is the library's fixed per-GPU array)."""
    uu, vv = _f64(u), _f64(v)
    tp = _f64(taper) if taper is not None else None
    V = np.zeros((frames, len(uu), 4, 2))
    fn = cache._lib.ks_visibilities_noalloc if noalloc else cache._lib.ks_visibilities
    _check(fn(cache._h, _i64(frames), _d(psize), _ptr(uu), _ptr(vv), _ptr(tp), _i(len(uu)), _ptr(V)))
    return V[..., 0] + 1j * V[..., 1]


def gradient_forward(cache, model, times, nus, L, nmax=-1, slab=0.0, copy=True):
    """The forward pass of sweep_passes(:dual) (Fit.jl:451-474): the transfer with the records the
    reverse pass needs; the observed images stay on the device. Returns them in sorted order,
    (nnu, nt, npix, 4), when copy, else None."""
    if model.kind == "thin":
        raise ValueError("the dual sweep takes polarized splats (thermal, power-law, kappa, knot or composite)")
    t, n = _f64(np.atleast_1d(times)), _f64(np.atleast_1d(nus))
    out = np.zeros((len(n), len(t), cache.npix, 4)) if copy else None
    _check(cache._lib.ks_gradient_forward(cache._h, model._h, _ptr(t), _i(len(t)), _ptr(n), _i(len(n)), _d(L), _i(nmax), _d(slab),
                                          _ptr(out), _i(0)))
    cache._fw = (len(t), len(n))
    return out


def gradient_reverse(cache, model, dparams, dstokes=None):
    """The reverse pass of sweep_passes(:dual): adds d(sum dstokes . I_obs)/dparams of the last
    forward pass into dparams (Fortran-ordered, shaped like the parameters). dstokes in sorted order,
    (nnu, nt, npix, 4); None takes the device seeds that frame_chi2_seed / image_chi2_seed wrote."""
    if dparams.shape != model.grad_shape or not dparams.flags.f_contiguous or dparams.dtype != np.float64:
        raise DimensionMismatch("dparams must be a Fortran-ordered float64 array shaped like the parameters (a vector for composites)")
    ds = None
    if dstokes is not None:
        nt, nnu = getattr(cache, "_fw", (0, 0))
        ds = np.ascontiguousarray(np.asarray(dstokes, dtype=np.float64))
        if ds.size != nnu * nt * cache.npix * 4:
            raise DimensionMismatch(f"dstokes must hold {(nnu, nt, cache.npix, 4)} values, got {ds.shape}")
    _check(cache._lib.ks_gradient_reverse(cache._h, model._h, _ptr(ds), _i(0), _ptr(dparams)))
    return dparams


def vis_chi2_seed(cache, frame, psize, u, v, vis, sigma, taper=None, gains=None, return_model=False):
    """frame_chi2_seed! (device_visibilities.jl:97-111) for frame `frame` (it + nt inu) of the last
    forward pass: chi2 = sum |(g V - d) / sigma|^2 over baselines and Stokes parameters, and the
    adjoint seed of the weighted residuals left on the device for gradient_reverse(dstokes=None).
    vis: complex (nuv, 4) [Jy]; sigma: (nuv, 4) or a 4-vector [Jy]; gains: complex (nuv,) or None."""
    uu, vv = _f64(u), _f64(v)
    nuv = len(uu)
    d = np.asarray(vis, dtype=np.complex128)
    if d.shape != (nuv, 4):
        raise DimensionMismatch(f"vis must be ({nuv}, 4), got {d.shape}")
    dv = np.ascontiguousarray(np.stack([d.real, d.imag], axis=-1))
    sg = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (nuv, 4)).copy()
    tp = _f64(taper) if taper is not None else None
    g = None
    if gains is not None:
        gc = np.asarray(gains, dtype=np.complex128)
        g = np.ascontiguousarray(np.stack([gc.real, gc.imag], axis=-1))
    chi = np.zeros(1)
    V = np.zeros((nuv, 4, 2)) if return_model else None
    _check(cache._lib.ks_vis_chi2_seed(cache._h, _i64(frame), _d(psize), _ptr(uu), _ptr(vv), _ptr(tp), _ptr(g), _ptr(dv), _ptr(sg),
                                       _i(nuv), _ptr(chi), _ptr(V)))
    if return_model:
        return float(chi[0]), V[..., 0] + 1j * V[..., 1]
    return float(chi[0])


def closure_chi2_seed(cache, frame, psize, data, taper=None, add=False):
    """chi2_closures (visibilities.jl:141-199) of frame `frame` of the last forward pass on the device: the closure
    phases and log closure amplitudes of the model's Stokes I at data.u, data.v against `data` (a ClosureData:
    1-based signed baseline indices, phases and sigmas, log amplitudes and sigmas), the chi2 returned and its
    adjoint seed written on the device (added to the frame's seed with add=True)."""
    uu, vv = _f64(data.u), _f64(data.v)
    nuv = len(uu)
    tri = np.ascontiguousarray(np.asarray(data.triangles, dtype=np.int32).reshape(-1, 3))
    quad = np.ascontiguousarray(np.asarray(data.quadrangles, dtype=np.int32).reshape(-1, 4))
    ntri, nquad = len(tri), len(quad)
    cp, sp = _f64(data.phases), _f64(data.sigma_phase)
    la, sl = _f64(data.logamps), _f64(data.sigma_logamp)
    if len(cp) != ntri or len(sp) != ntri or len(la) != nquad or len(sl) != nquad:
        raise DimensionMismatch("closure data: one phase and sigma per triangle, one log amplitude and sigma per quadrangle")
    tp = _f64(taper) if taper is not None else None
    chi = np.zeros(1)
    ip = ctypes.POINTER(ctypes.c_int)
    _check(cache._lib.ks_closure_chi2_seed(cache._h, _i64(frame), _d(psize), _ptr(uu), _ptr(vv), _ptr(tp), _i(nuv), _i(ntri),
                                           tri.ctypes.data_as(ip) if ntri else None, _ptr(cp) if ntri else None,
                                           _ptr(sp) if ntri else None, _i(nquad), quad.ctypes.data_as(ip) if nquad else None,
                                           _ptr(la) if nquad else None, _ptr(sl) if nquad else None, _i(int(add)), _ptr(chi)))
    return float(chi[0])


def image_chi2_seed(cache, frame, data, sigma, mask=None, binning=None):
    """_frame_seed! (Fit.jl:537-577) for frame `frame` of the last forward pass: chi2 = sum ((I - d) /
    sigma)^2 over the masked pixels and the seed 2 r / sigma on the device; with a Binning, I is the
    binned image and each point receives its pixel's seed divided by the pixel's point count.
    data: (n_alpha, n_beta, 4) observed Stokes [cgs] of the (binned) screen; sigma: a 4-vector or
    (n_alpha, n_beta, 4); mask: (n_alpha, n_beta) booleans or None."""
    na, nb = cache.screen_size if binning is None else binning.size
    dd = np.asarray(data, dtype=np.float64)
    if dd.shape != (na, nb, 4):
        raise DimensionMismatch(f"data must be {(na, nb, 4)}, got {dd.shape}")
    flat = lambda x: np.ascontiguousarray(np.transpose(x, (1, 0, 2)).reshape(na * nb, 4))   # pixel = i + na j
    sg = np.asarray(sigma, dtype=np.float64)
    per_pixel = sg.shape != (4,)
    sgf = flat(np.broadcast_to(sg, (na, nb, 4))) if per_pixel else np.ascontiguousarray(sg)
    mk = None
    if mask is not None:
        mm = np.asarray(mask, dtype=bool)
        if mm.shape != (na, nb):
            raise DimensionMismatch(f"mask must be {(na, nb)}, got {mm.shape}")
        mk = np.ascontiguousarray(mm.T.reshape(na * nb).astype(np.uint8))
    chi = np.zeros(1)
    if binning is None:
        _check(cache._lib.ks_image_chi2_seed(cache._h, _i64(frame), _ptr(flat(dd)), _ptr(sgf), _i(int(per_pixel)), _ptr(mk), _ptr(chi)))
    else:
        _check(cache._lib.ks_image_chi2_seed_binned(cache._h, _i64(frame), binning.handle(), _ptr(flat(dd)), _ptr(sgf), _i(int(per_pixel)),
                                                    _ptr(mk), _ptr(chi)))
    return float(chi[0])


def vis_seed(cache, frame, psize, u, v, w, taper=None, add=False):
    """seed_kernel! (device_visibilities.jl:76-88) for any visibility-domain loss: the adjoint transform of
    w = dl/dRe V + i dl/dIm V (complex (nuv, 4)) into the device seed of frame `frame` of the last
    forward pass (added to the existing seed with add=True), for gradient_reverse(dstokes=None)."""
    uu, vv = _f64(u), _f64(v)
    ww = np.asarray(w, dtype=np.complex128)
    if ww.shape != (len(uu), 4):
        raise DimensionMismatch(f"w must be ({len(uu)}, 4), got {ww.shape}")
    wv = np.ascontiguousarray(np.stack([ww.real, ww.imag], axis=-1))
    tp = _f64(taper) if taper is not None else None
    _check(cache._lib.ks_vis_seed(cache._h, _i64(frame), _d(psize), _ptr(uu), _ptr(vv), _ptr(tp), _ptr(wv), _i(len(uu)), _i(int(add))))
