"""ctypes bindings of the CPU oracle (oracle/build/libks_oracle.so).

TEST INFRASTRUCTURE. The oracle is the FP64 CPU restatement of KerrSplat's geodesic march and
polarized transfer that the HIP product is checked against. Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg may import this module; the product path never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libks_oracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.kso_planck.restype = _d
        _lib.kso_thermal_leung.restype = _d
        _lib.kso_kappa_hyp.restype = _d
        for name in ("kso_planck", "kso_thermal_leung", "kso_kappa_hyp"):
            getattr(_lib, name).argtypes = None
    return _lib


def _p(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().kso_set_threads(_i(n))


def num_threads():
    return lib().kso_num_threads()


def pixel_constants(a, thetao, alpha, beta):
    out = np.zeros(17)
    lib().kso_pixel_constants(_d(a), _d(thetao), _d(alpha), _d(beta), _p(out))
    return {"nreal": int(out[0]), "case": int(out[1]), "tau_total": out[2], "fo": out[3], "k": out[4],
            "lam": out[5], "eta": out[6], "up": out[7], "um": out[8],
            "roots": out[9:17].reshape(4, 2)}


def root_cases(a, thetao, alphas, betas):
    al, be = _f64(alphas), _f64(betas)
    out = np.zeros(len(al), dtype=np.int8)
    lib().kso_root_cases(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _p(out))
    return out


def case_permutation(numreals):
    nr = np.ascontiguousarray(numreals, dtype=np.int8)
    perm = np.zeros(len(nr), dtype=np.int64)
    counts = np.zeros(3, dtype=np.int64)
    lib().kso_case_permutation(_i(len(nr)), _p(nr), _p(perm), _p(counts))
    return perm, counts


def samples(a, thetao, alphas, betas, N, M=64, marcher=1):
    al, be = _f64(alphas), _f64(betas)
    n = len(al)
    t, r, th, ph = (np.zeros((n, N)) for _ in range(4))
    flags = np.zeros((n, N), dtype=np.uint8)
    rt, rp, tt = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().kso_samples(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(marcher),
                      _p(t), _p(r), _p(th), _p(ph), _p(flags), _p(rt), _p(rp), _p(tt))
    return {"t": t, "r": r, "theta": th, "phi": ph, "flags": flags, "resid_t": rt, "resid_phi": rp, "tau_total": tt}


def ray_hits(a, thetao, alphas, betas, N, params, times, nmax=-1, slab=0.0, M=64, marcher=1):
    """ray_lists' hit table (dual_gradient.jl:283-324): hits[i, j, f] = 1 when a step-taking sample of ray j
    (screen order; before the half-orbit cutoff when nmax >= 0) enters splat i's support sphere at times[f]."""
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    tv = _f64(np.atleast_1d(times))
    ns, n, nt = P.shape[1], len(al), len(tv)
    hits = np.zeros(ns * n * nt, dtype=np.uint8)
    lib().kso_ray_hits(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(marcher), _i(P.shape[0]), _i(ns), _p(P), _i(nt), _p(tv),
                       _i(nmax), _d(slab), _p(hits))
    return hits.reshape(nt, n, ns).transpose(2, 1, 0)


def thin_image(a, thetao, alphas, betas, N, params, t_obs, M=64, marcher=1):
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    out = np.zeros(len(al))
    lib().kso_thin_image(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(marcher),
                         _i(P.shape[1]), _p(P), _d(t_obs), _p(out))
    return out


KIND = {"thin": 0, "thermal": 1, "powerlaw": 2, "kappa": 3, "knot": 4}


def polarized_image(a, thetao, alphas, betas, N, params, t_obs, nu, L, kind="thermal", hyp=None, gamma_max=1e5,
                    knots=None, knot_times=None, nmax=-1, slab=0.0, M=64, marcher=1):
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    hy = _f64(hyp) if hyp is not None else None
    kn = np.asfortranarray(knots, dtype=np.float64) if knots is not None else None
    kt = _f64(knot_times) if knot_times is not None else None
    nk = 0 if kt is None else len(kt)
    n = len(al)
    out = np.zeros((n, 4))
    rt, rp = np.zeros(n), np.zeros(n)
    lib().kso_polarized_image(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(marcher),
                              _i(KIND[kind]), _i(P.shape[0]), _i(P.shape[1]), _p(P), _p(hy), _d(gamma_max),
                              _p(kn), _p(kt), _i(nk), _d(t_obs), _d(nu), _d(L), _i(nmax), _d(slab),
                              _p(out), _p(rt), _p(rp))
    return out, rt, rp


def spacetime_tangent(a, thetao, alphas, betas, N, params, t_obs, nu, L, kind="thermal", hyp=None, gamma_max=1e5, nmax=-1,
                      slab=0.0, M=64, marcher=1):
    """Observed Stokes (npix, 4) and their forward-mode partials (npix, 4, 3) with respect to (a, theta_o, ln L)
    (spacetime_valgrad / spacetime_jacobian, joint.jl:74-154)."""
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    hy = _f64(hyp) if hyp is not None else None
    n = len(al)
    out, dout = np.zeros((n, 4)), np.zeros((n, 4, 3))
    rc = lib().kso_spacetime_tangent(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(marcher), _i(KIND[kind]),
                                     _i(P.shape[0]), _i(P.shape[1]), _p(P), _p(hy), _d(gamma_max), _d(t_obs), _d(nu),
                                     _d(np.log(L)), _i(nmax), _d(slab), _p(out), _p(dout))
    if rc:
        raise ValueError(f"kso_spacetime_tangent: unsupported model (code {rc})")
    return out, dout


def _parts(parts):
    """ctypes arrays of composite parts: (kind, params[, hyp[, gamma_max[, knots, knot_times]]]) for thermal,
    powerlaw or kappa parts, ("knot", params, knots, knot_times) for thermal knot parts; knots 3 x nknots x ns."""
    k = len(parts)
    kinds, rows, ns, nk = (ctypes.c_int * k)(), (ctypes.c_int * k)(), (ctypes.c_int * k)(), (ctypes.c_int * k)()
    Ps, Hs, gm = (ctypes.c_void_p * k)(), (ctypes.c_void_p * k)(), (ctypes.c_double * k)()
    Ks, Ts = (ctypes.c_void_p * k)(), (ctypes.c_void_p * k)()
    keep = []
    for q, part in enumerate(parts):
        kind, params = part[0], part[1]
        P = np.asfortranarray(params, dtype=np.float64)
        kinds[q], rows[q], ns[q] = KIND[kind], P.shape[0], P.shape[1]
        Ps[q] = P.ctypes.data
        gm[q] = 1e5
        keep.append(P)
        if kind == "knot":
            kn_, kt_ = part[2], part[3]
        else:
            hyp = part[2] if len(part) > 2 else None
            hy = _f64(hyp) if hyp is not None else None
            keep.append(hy)
            Hs[q] = hy.ctypes.data if hy is not None else None
            gm[q] = part[3] if len(part) > 3 and part[3] is not None else 1e5
            kn_, kt_ = (part[4], part[5]) if len(part) > 5 else (None, None)
        if kn_ is not None:
            kn = np.asfortranarray(kn_, dtype=np.float64)
            kt = _f64(kt_)
            keep += [kn, kt]
            Ks[q], Ts[q], nk[q] = kn.ctypes.data, kt.ctypes.data, len(kt)
    return (kinds, rows, ns, Ps, Hs, gm, Ks, Ts, nk), keep


def polarized_composite(a, thetao, alphas, betas, N, parts, t_obs, nu, L, nmax=-1, slab=0.0, M=64, marcher=1):
    """Stokes image of CompositeModel(parts...) (transport.jl:298-316). parts: a list of
    (kind, params[, hyp[, gamma_max[, knots, knot_times]]]) with kind thermal, powerlaw or kappa, or
    ("knot", params, knots, knot_times) with knots 3 x nknots x ns."""
    al, be = _f64(alphas), _f64(betas)
    n = len(al)
    arrs, keep = _parts(parts)
    out = np.zeros((n, 4))
    rt, rp = np.zeros(n), np.zeros(n)
    rc = lib().kso_polarized_composite(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(marcher), _i(len(parts)), *arrs,
                                       _d(t_obs), _d(nu), _d(L), _i(nmax), _d(slab), _p(out), _p(rt), _p(rp))
    if rc:
        raise ValueError(f"kso_polarized_composite: bad parts (code {rc})")
    return out, rt, rp


def model_gradient(a, thetao, alphas, betas, N, parts, t_obs, nu, L, dstokes, nmax=-1, slab=0.0, M=64):
    """The dual-sweep gradient of sum dstokes . image for any model (parts as in polarized_composite; a single
    population is one part), every row of every splat by the generic element adjoint: returns (a list of
    nrows x ns gradient blocks, one per part, and the image npix x 4)."""
    al, be = _f64(alphas), _f64(betas)
    n = len(al)
    arrs, keep = _parts(parts)
    ds = _f64(dstokes)
    sizes = [np.asarray(p[1]).shape for p in parts]
    g = np.zeros(sum(r * c for r, c in sizes))
    img = np.zeros((n, 4))
    rc = lib().kso_model_gradient(_d(a), _d(thetao), _i(n), _p(al), _p(be), _i(N), _i(M), _i(len(parts)), *arrs,
                                  _d(t_obs), _d(nu), _d(L), _i(nmax), _d(slab), _p(ds), _p(g), _p(img))
    if rc:
        raise ValueError(f"kso_model_gradient: bad parts (code {rc})")
    out, off = [], 0
    for r, c in sizes:
        out.append(g[off:off + r * c].reshape((r, c), order="F").copy())
        off += r * c
    return out, img


def count_polarized(a, thetao, alphas, betas, N, params, t_obs, nu, L, kind="thermal", hyp=None, gamma_max=1e5, M=64):
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    hy = _f64(hyp) if hyp is not None else None
    f, s = _d(0), _d(0)
    lib().kso_count_polarized(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(KIND[kind]),
                              _i(P.shape[0]), _i(P.shape[1]), _p(P), _p(hy), _d(gamma_max), _d(t_obs), _d(nu), _d(L),
                              ctypes.byref(f), ctypes.byref(s))
    return f.value, s.value


def thermal(ne, Th, B, nu, th, dexter_rhoV=False, pandya=False):
    out = np.zeros(8)
    lib().kso_thermal(_d(ne), _d(Th), _d(B), _d(nu), _d(th), _i(int(dexter_rhoV)), _i(int(pandya)), _p(out))
    return out


def thermal_pandya(ne, Th, B, nu, th):
    out = np.zeros(3)
    lib().kso_thermal_pandya(_d(ne), _d(Th), _d(B), _d(nu), _d(th), _p(out))
    return out


def thermal_leung(ne, Th, B, nu, th):
    f = lib().kso_thermal_leung
    f.argtypes = [_d] * 5
    return f(ne, Th, B, nu, th)


def powerlaw(ne, p, gmin, gmax, B, nu, th, rotativities=True):
    out = np.zeros(8)
    lib().kso_powerlaw(_d(ne), _d(p), _d(gmin), _d(gmax), _d(B), _d(nu), _d(th), _i(int(rotativities)), _p(out))
    return out


def kappa(ne, kap, w, B, nu, th, hyp):
    out = np.zeros(8)
    lib().kso_kappa(_d(ne), _d(kap), _d(w), _d(B), _d(nu), _d(th), _d(hyp), _p(out))
    return out


def kappa_hyp(kap, w):
    f = lib().kso_kappa_hyp
    f.argtypes = [_d, _d]
    return f(kap, w)


def besselk01(x):
    k0, k1 = _d(0), _d(0)
    lib().kso_besselk01(_d(x), ctypes.byref(k0), ctypes.byref(k1))
    return k0.value, k1.value


def planck(nu, Th):
    f = lib().kso_planck
    f.argtypes = [_d, _d]
    return f(nu, Th)


def step_operator(alpha4, rho3, D):
    al, rh = _f64(alpha4), _f64(rho3)
    O, E = np.zeros(16), np.zeros(16)
    lib().kso_step_operator(_p(al), _p(rh), _d(D), _p(O), _p(E))
    return O.reshape(4, 4), E.reshape(4, 4)


def cap(c8):
    c = _f64(c8)
    out = np.zeros(8)
    lib().kso_cap(_p(c), _p(out))
    return out


def local_frame(a, r, th, eta, lam, nur, nuth, alpha, beta, thetao, u, B):
    uu, BB = _f64(u), _f64(B)
    out = np.zeros(3)
    lib().kso_local_frame(_d(a), _d(r), _d(th), _d(eta), _d(lam), _i(int(nur)), _i(int(nuth)), _d(alpha), _d(beta),
                          _d(thetao), _p(uu), _p(BB), _p(out))
    return out


def elliptic(phi, m, n):
    out = np.zeros(7)
    lib().kso_elliptic(_d(phi), _d(m), _d(n), _p(out))
    return {"F": out[0], "E": out[1], "Pi": out[2], "sn": out[3], "cn": out[4], "dn": out[5], "am": out[6]}


def visibilities(images, nalpha, nbeta, psize, u, v, taper=None):
    """images: (nframes, npix, 4) screen order; returns complex (nframes, nuv, 4)."""
    im = _f64(images)
    nf = im.shape[0]
    uu, vv = _f64(u), _f64(v)
    tp = _f64(taper) if taper is not None else None
    Vr, Vi = np.zeros((nf, len(uu), 4)), np.zeros((nf, len(uu), 4))
    lib().kso_visibilities(_i(nalpha), _i(nbeta), _i(nf), _p(im), _d(psize), _i(len(uu)), _p(uu), _p(vv), _p(tp), _p(Vr), _p(Vi))
    return Vr + 1j * Vi


def vis_seed(W, nalpha, nbeta, psize, u, v, taper=None):
    Wc = np.asarray(W)
    nf = Wc.shape[0]
    Wr, Wi = _f64(Wc.real), _f64(Wc.imag)
    uu, vv = _f64(u), _f64(v)
    tp = _f64(taper) if taper is not None else None
    seed = np.zeros((nf, nalpha * nbeta, 4))
    lib().kso_vis_seed(_i(nalpha), _i(nbeta), _i(nf), _d(psize), _i(len(uu)), _p(uu), _p(vv), _p(tp), _p(Wr), _p(Wi), _p(seed))
    return seed


def count_gradient(a, thetao, alphas, betas, N, params, t_obs, nu, L, dstokes, M=64):
    """Operation counts of the dual sweep on the Counted scalar, per phase (march, forward steps,
    tails, adjoint sweep, and the sweep's per-sample step adjoints): returns (flops[5], special[5])
    summed over the rays."""
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    ds = _f64(dstokes)
    c, sp = np.zeros(5), np.zeros(5)
    lib().kso_count_gradient(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(P.shape[1]), _p(P),
                             _d(t_obs), _d(nu), _d(L), _p(ds), _p(c), _p(sp))
    return c, sp


def polarized_gradient(a, thetao, alphas, betas, N, params, t_obs, nu, L, dstokes, nmax=-1, slab=0.0, M=64, analytic_weight=False):
    """The dual-sweep gradient (thermal splats): returns (dparams 21 x ns, image npix x 4). The splat
    weight's derivative by 13-partial duals, or written out by hand with analytic_weight."""
    al, be = _f64(alphas), _f64(betas)
    P = np.asfortranarray(params, dtype=np.float64)
    ds = _f64(dstokes)
    g = np.zeros(P.shape, order="F")
    img = np.zeros((len(al), 4))
    lib().kso_polarized_gradient(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(P.shape[1]), _p(P),
                                 _d(t_obs), _d(nu), _d(L), _i(nmax), _d(slab), _p(ds), g.ctypes.data_as(_P), _p(img),
                                 _i(int(analytic_weight)))
    return g, img


# ---- test-only fluid models (ks_fluid.hpp): end-to-end pins of the transport -------------------------------
def riaf_image(a, thetao, alphas, betas, N, nu, L, pandya=True, M=64, marcher=1):
    """Observed Stokes (npix, 4) of ipole's example RIAF (the reference's test/riaf_model.jl)."""
    al, be = _f64(alphas), _f64(betas)
    out = np.zeros((len(al), 4))
    lib().kso_riaf_image(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(marcher), _d(nu), _d(L),
                         _i(int(pandya)), _p(out))
    return out


def slab_image(a, thetao, alphas, betas, N, nu, L, slabs, M=64, marcher=1):
    """Observed Stokes (npix, 4) of shells with given fluid-frame coefficients: slabs is a list of
    (r1, r2, (jI, jQ, jV, aI, aQ, aV, rQ, rV), u[3], B[3])."""
    al, be = _f64(alphas), _f64(betas)
    rows = [np.concatenate([[r1, r2], c, u, B]) for (r1, r2, c, u, B) in slabs]
    sl = _f64(np.array(rows).reshape(-1, 16))
    out = np.zeros((len(al), 4))
    lib().kso_slab_image(_d(a), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(marcher), _d(nu), _d(L),
                         _i(len(rows)), _p(sl), _p(out))
    return out


def gold_image(model, thetao, alphas, betas, N, nu, L, polarized=False, M=64, marcher=1):
    """Gold et al. (2020) model `model` (1..5): (spin, observed intensity (npix,)), or with polarized the
    Stokes (npix, 4) of the same model through the polarized transport."""
    al, be = _f64(alphas), _f64(betas)
    out = np.zeros((len(al), 4)) if polarized else np.zeros(len(al))
    spin = _d(0.0)
    rc = lib().kso_gold_image(_i(model - 1), _d(thetao), _i(len(al)), _p(al), _p(be), _i(N), _i(M), _i(marcher), _d(nu), _d(L),
                              _i(int(polarized)), _p(out), ctypes.byref(spin))
    if rc:
        raise ValueError(f"no Gold model {model}")
    return spin.value, out
