"""End-to-end pins of the oracle's transport against independent codes and exact symmetries.

* Gold et al. (2020, ApJ 897, 148) GRRT test problems 1-5: total fluxes at 128^2 x 2000 samples within 2% of
  the published values (the reference's gate, test/test_gold2020.jl:55-85). Pins the geodesics, the sample
  grid, the redshift, the unpolarized transfer and the unit conversions together.
* The Gold torus through the polarized transport equals the unpolarized transport (test_polarized.jl:66-82).
* ipole's RIAF: full Stokes images at 230 GHz and 2 THz against ipole's own (tests/golden/ipole_riaf, the
  reference's validation/ipole_riaf CSVs) at the reference's bars (test_polarized.jl:156-181): pixel-norm
  residuals < 3%, totals < 1%, median EVPA difference < 0.01 rad, linear-polarization fraction < 0.005. A
  moving, plunging fluid with a toroidal field: pins the ZAMO-to-fluid boost, the pitch angle, the
  Walker-Penrose screen angle, the thermal coefficients (Pandya emissivities, Dexter rho_Q, Shcherbakov
  rho_V) and the 4x4 step with Faraday depths up to ~9 rad.
* A Faraday screen rotates the EVPA by (1/2) sum rho_V x (fluid path length) computed from the samples, keeps
  I and P, and two co-located screens add (test_polarized.jl:84-119); co-located emitters add and absorbing
  elements commute (test_polarized.jl:121-134).
* Equatorial and azimuthal reflections (test_polarized_splats.jl:197-241): the image at theta_o = 163 deg is
  the beta-mirror of the reflected source at 17 deg, the image at spin -a the alpha-mirror at +a, with U and V
  reversed; a polar-vector mirror of the field breaks it.
* Knot splats: constant knots at the centre render as the same splat without pattern rotation (to rounding), and
  knots sampled densely along the rotating centre converge to the rotating splat at second order in the knot
  spacing (linear interpolation).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

MSUN, GNEWT, CL, PC, JY = 1.989e33, 6.6742e-8, 2.99792458e10, 3.085678e18, 1e-23


def grav(m):
    return GNEWT * m * MSUN / CL ** 2


def flat(xs, ys):
    """Screen-order (column-major, alpha fastest) pixel coordinates of the grid xs x ys."""
    return np.repeat(xs[:, None], len(ys), 1).ravel(order="F"), np.repeat(ys[None, :], len(xs), 0).ravel(order="F")


GOLD_PUBLISHED = (1.6465, 1.4360, 0.4418, 0.2710, 0.0255)


def gold_camera(res=128, fov=30.0):
    d = fov / res
    ax = np.array([-fov / 2 + (k + 0.5) * d for k in range(res)])
    return flat(ax, ax), d


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_gold2020_fluxes(oracle, model):
    (al, be), d = gold_camera()
    L = grav(4.063e6)
    a, I = oracle.gold_image(model, math.radians(60.0), al, be, 2000, 230e9, L)
    F = I.sum() * (d * L / (7778.0 * PC)) ** 2 / JY
    assert np.all(np.isfinite(I)) and np.all(I >= 0)
    assert abs(F / GOLD_PUBLISHED[model - 1] - 1) < 0.02, (model, F)


def test_gold_through_the_polarized_path(oracle):
    (al, be), _ = gold_camera(24)
    L = grav(4.063e6)
    _, I = oracle.gold_image(4, math.radians(60.0), al, be, 400, 230e9, L)
    _, S = oracle.gold_image(4, math.radians(60.0), al, be, 400, 230e9, L, polarized=True)
    assert I.max() > 0
    assert np.abs(S[:, 0] - I).max() <= 1e-12 * I.max()
    assert np.all(S[:, 1:] == 0.0)


def riaf_camera(res=100):
    L = grav(4.3e6)
    D = 8.3e3 * PC
    fov = 200e-6 / 206264.806247 * D / L
    xs = ((np.arange(res) + 0.49) / res - 0.5) * fov        # ipole's pixel centres (model_geodesics.c init_XK)
    ys = ((np.arange(res) + 0.5) / res - 0.5) * fov
    return flat(xs, ys), (fov / res * L / D) ** 2 / JY, L


@pytest.mark.parametrize("tag,nu", [("riaf_fine", 230e9), ("riaf_2thz", 2e12)])
def test_riaf_vs_ipole(oracle, tag, nu):
    res = 100
    (al, be), scale, L = riaf_camera(res)
    S = oracle.riaf_image(0.9375, math.radians(85.0), al, be, 2000, nu, L)
    ours = [S[:, k].reshape(res, res, order="F") for k in range(4)]
    ref = [np.loadtxt(os.path.join(GOLDEN, "ipole_riaf", f"{tag}_{s}.csv"), delimiter=",") for s in "IQUV"]
    resid = [np.linalg.norm(ours[k] - ref[k]) / np.linalg.norm(ref[k]) for k in range(4)]
    tot = np.array([ours[k].sum() * scale for k in range(4)])
    totr = np.array([ref[k].sum() * scale for k in range(4)])
    w = ref[0] > 0.01 * ref[0].max()
    dchi = np.abs(np.remainder(0.5 * np.arctan2(ours[2], ours[1]) - 0.5 * np.arctan2(ref[2], ref[1]) + np.pi / 2, np.pi) - np.pi / 2)
    lp = np.median(np.hypot(ours[1], ours[2])[w] / ours[0][w])
    lpr = np.median(np.hypot(ref[1], ref[2])[w] / ref[0][w])
    assert max(resid) < 0.03, resid
    assert np.all(np.abs(tot / totr - 1) < 0.01), (tot, totr)
    assert np.median(dchi[w]) < 0.01
    assert abs(lp - lpr) < 0.005


def faraday_slabs(rhoV):
    B = np.array([0.3, 0.8, 0.5])
    u0 = np.zeros(3)
    emit = (12.0, 14.0, np.array([1e-14, 0.6e-14, 0.1e-14, 0, 0, 0, 0, 0.0]), u0, B)
    screen = (16.0, 18.0, np.array([0, 0, 0, 0, 0, 0, 0, rhoV]), u0, B)
    return emit, screen


def test_faraday_screen_rotation(oracle):
    a, th, nu, L, N = 0.5, math.radians(60.0), 230e9, 1e12, 1500
    al, be = np.array([2.0]), np.array([1.0])
    rhoV = 2e-12
    emit, screen = faraday_slabs(rhoV)
    S0 = oracle.slab_image(a, th, al, be, N, nu, L, [emit])[0]
    S1 = oracle.slab_image(a, th, al, be, N, nu, L, [emit, screen])[0]
    S2 = oracle.slab_image(a, th, al, be, N, nu, L, [emit, screen, screen])[0]
    # rotation measure from the samples: nu_f rho_V x (L / nu_obs) Sigma dtau = rho_V x the fluid path length
    smp = oracle.samples(a, th, al, be, N)
    r, thk, fl = smp["r"], smp["theta"], smp["flags"]
    pc = oracle.pixel_constants(a, th, 2.0, 1.0)
    dtau = smp["tau_total"][0] / (N + 1)
    rm = 0.0
    for k in range(N):
        if not (fl[0, k] & 1) or not (16.0 <= r[0, k] <= 18.0):
            continue
        g, _, _ = oracle.local_frame(a, r[0, k], thk[0, k], pc["eta"], pc["lam"], bool(fl[0, k] & 2), bool(fl[0, k] & 4), 2.0, 1.0,
                                     th, np.zeros(3), screen[4])
        sig = r[0, k] ** 2 + a * a * math.cos(thk[0, k]) ** 2
        rm += nu / g * rhoV * L / nu * sig * dtau
    assert rm > 0.1
    chi = [0.5 * math.atan2(S[2], S[1]) for S in (S0, S1, S2)]
    P0, P1 = math.hypot(S0[1], S0[2]), math.hypot(S1[1], S1[2])
    assert abs(S1[0] / S0[0] - 1) < 1e-12
    assert abs(P1 / P0 - 1) < 1e-10
    assert abs(math.remainder(chi[1] - chi[0] - rm / 2, math.pi)) < 1e-9 * max(1, rm)
    assert abs(math.remainder(chi[2] - chi[0] - rm, math.pi)) < 1e-9 * max(1, rm)


def test_colocated_elements_add_and_commute(oracle):
    a, th, nu, L = 0.5, math.radians(60.0), 230e9, 1e12
    al, be = np.array([2.0]), np.array([1.0])
    e1 = (10.0, 13.0, np.array([1e-14, 0.5e-14, 0.2e-14, 0, 0, 0, 0, 0]), np.array([0.3, 0.4, 0.0]), np.array([1.0, 0.0, 0.2]))
    e2 = (10.0, 13.0, np.array([2e-14, -0.4e-14, 0.3e-14, 0, 0, 0, 0, 0]), np.array([-0.2, 0.5, 0.1]), np.array([0.0, 1.0, -0.3]))
    Sa = oracle.slab_image(a, th, al, be, 800, nu, L, [e1])[0]
    Sb = oracle.slab_image(a, th, al, be, 800, nu, L, [e2])[0]
    Sab = oracle.slab_image(a, th, al, be, 800, nu, L, [e1, e2])[0]
    assert np.abs(Sab - (Sa + Sb)).max() <= 1e-12 * np.abs(Sa + Sb).max()
    e3 = (10.0, 13.0, np.array([1e-14, 0.5e-14, 0.2e-14, 3e-13, 1e-13, 0.5e-13, 2e-13, 4e-13]), np.array([0.3, 0.4, 0.0]),
          np.array([1.0, 0.0, 0.2]))
    S12 = oracle.slab_image(a, th, al, be, 800, nu, L, [e3, e2])[0]
    S21 = oracle.slab_image(a, th, al, be, 800, nu, L, [e2, e3])[0]
    assert np.abs(S12 - S21).max() <= 1e-13 * np.abs(S12).max()


def reflection_params():
    p = np.zeros((21, 2))
    p[:, 0] = [5.0, 2.0, 0.5, math.log(1.5), math.log(1.5), math.log(1.0), 1.0, 0, 0, 0, 0.0, math.log(1e9), math.log(3e6),
               math.log(30.0), math.log(8.0), 1.0, 0.5, 0.3, 0.3, 0.1, 0.0]
    p[:, 1] = [-3.0, -4.0, -1.0, math.log(1.0), math.log(1.2), math.log(0.8), 1.0, 0, 0, 0, 0.0, math.log(1e9), math.log(1e5),
               math.log(20.0), math.log(15.0), 2.0, -1.0, -0.2, 0.4, -0.1, 0.0]
    return p


def reflection_image(oracle, a, th, P, res=24, N=60):
    fov = 20.0
    d = fov / res
    ax = np.linspace(-fov / 2 + d / 2, fov / 2 - d / 2, res)
    al, be = flat(ax, ax)
    S, _, _ = oracle.polarized_image(a, th, al, be, N, P, 0.0, 230e9, grav(6.5e9))
    return S.reshape(res, res, 4, order="F")


def rel(x, y):
    return math.sqrt(((x - y) ** 2).sum() / (y ** 2).sum())


def test_reflection_parity(oracle):
    p = reflection_params()
    A = reflection_image(oracle, 0.94, math.radians(163.0), p)
    q = p.copy(); q[2] = -q[2]; q[16] += math.pi; q[19] = -q[19]          # z -> -z: B -> -(mirror), u_z reversed
    D = reflection_image(oracle, 0.94, math.radians(17.0), q)
    sgn = np.array([1, 1, -1, -1])
    for k in range(4):
        assert rel(D[:, ::-1, k] * sgn[k], A[:, :, k]) < 1e-10, k
    B = reflection_image(oracle, 0.94, math.radians(17.0), p)
    r = p.copy(); r[1] = -r[1]; r[15] = math.pi - r[15]; r[16] = math.pi - r[16]; r[18] = -r[18]   # y -> -y, spin reversed
    E = reflection_image(oracle, -0.94, math.radians(17.0), r)
    for k in range(4):
        assert rel(E[::-1, :, k] * sgn[k], B[:, :, k]) < 1e-10, k
    q2 = q.copy(); q2[16] -= math.pi; q2[15] = math.pi - q2[15]            # a polar-vector mirror of the field
    D2 = reflection_image(oracle, 0.94, math.radians(17.0), q2)
    assert rel(-D2[:, ::-1, 3], A[:, :, 3]) > 0.1


def knot_params(omega):
    p = np.zeros((21, 1))
    p[:, 0] = [6.0, 1.0, 0.3, math.log(1.2), math.log(1.2), math.log(1.2), 1.0, 0.2, -0.1, 0.3, 0.0, math.log(1e9), math.log(3e6),
               math.log(25.0), math.log(10.0), 1.1, 0.4, 0.2, 0.3, -0.1, omega]
    return p


def test_constant_knots_equal_the_splat_without_rotation(oracle):
    a, th, res, N = 0.5, math.radians(60.0), 16, 128
    ax = np.linspace(-12.0, 12.0, res)
    al, be = flat(ax, ax)
    P = knot_params(0.0)
    times = np.array([-200.0, -50.0, 10.0, 100.0])
    knots = np.repeat(P[0:3, 0][:, None], 4, axis=1)[:, :, None]
    L = grav(6.5e9)
    Sk, _, _ = oracle.polarized_image(a, th, al, be, N, P, 30.0, 230e9, L, kind="knot", knots=knots, knot_times=times)
    Ss, _, _ = oracle.polarized_image(a, th, al, be, N, P, 30.0, 230e9, L)
    assert np.abs(Ss[:, 0]).max() > 0
    # (1 - w) c + w c is c up to rounding
    assert np.abs(Sk - Ss).max() <= 1e-13 * np.abs(Ss[:, 0]).max()


def test_dense_knots_converge_to_the_rotating_splat(oracle):
    """An isotropic splat whose centre turns at omega about the spin axis is a knot splat with knots on the
    circle; linear interpolation between knots spaced h in time misses by O(h^2)."""
    a, th, res, N = 0.5, math.radians(60.0), 14, 128
    ax = np.linspace(-12.0, 12.0, res)
    al, be = flat(ax, ax)
    om = 0.05
    P = knot_params(om)
    P[6:10, 0] = [1.0, 0.0, 0.0, 0.0]           # isotropic widths: the pattern's orientation does not matter
    L = grav(6.5e9)
    t_obs = 30.0
    Ss, _, _ = oracle.polarized_image(a, th, al, be, N, P, t_obs, 230e9, L)
    Pk = P.copy(); Pk[20, 0] = 0.0
    errs = []
    for nk in (17, 33, 65):
        times = np.linspace(-150.0, 60.0, nk)
        ph = om * (times - P[10, 0])
        x0, y0, z0 = P[0:3, 0]
        knots = np.stack([x0 * np.cos(ph) - y0 * np.sin(ph), x0 * np.sin(ph) + y0 * np.cos(ph), np.full(nk, z0)])[:, :, None]
        Sk, _, _ = oracle.polarized_image(a, th, al, be, N, Pk, t_obs, 230e9, L, kind="knot", knots=knots, knot_times=times)
        errs.append(np.abs(Sk[:, 0] - Ss[:, 0]).max() / np.abs(Ss[:, 0]).max())
    assert errs[0] < 0.2
    for e0, e1 in zip(errs, errs[1:]):
        assert 3.0 < e0 / e1 < 5.5, errs
