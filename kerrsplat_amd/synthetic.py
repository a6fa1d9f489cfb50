"""Seeded synthetic inputs of the benchmark configurations (splat draws, uv coverage, physical scale).

The draws follow the configuration text of the benchmark (numpy PCG64 streams, seed
20261202 + config index):
  centres r ~ U[4, 10] M, theta ~ pi/2 + N(0, 0.15), phi ~ U[0, 2 pi) in Kerr-Schild Cartesian
  coordinates; widths log-U[0.5, 2] M per axis (rows s = ln w); quaternions normalized N(0, 1)^4;
  t0 = 0; logw = ln 1e9 for a single frame; pattern rate omega = 1/(r^{3/2} + a) (polarized.jl:224);
  thin: amplitudes log-U[0.1, 1]; thermal: n_e log-U[1e5, 1e7] cm^-3, Theta_e log-U[3, 50],
  |B| log-U[1, 30] G along a random direction, ZAMO velocity = Keplerian x U[0.7, 1]
  (keplerian_zamo_velocity, Fit.jl:787-796); power law: p ~ U[2, 3.5], ln gamma_min = ln 10;
  kappa: kappa ~ U[3, 7], w ~ log-U[3, 30]; uv points uniform in a disc of 10 G-lambda.
The length unit and distance are M = 6.5e9 M_sun and D = 16.8 Mpc (test_fit.jl:263).
"""
import math

import numpy as np

PC = 3.085678e18
MASS_MSUN = 6.5e9
DISTANCE_CM = 16.8e6 * PC


def seed(cfg):
    return np.random.Generator(np.random.PCG64(20261202 + cfg))


def quasi_cartesian_ks(a, r, th, ph):
    """BL -> quasi-Cartesian Kerr-Schild (coordinates.jl:22-31)."""
    temp = math.sqrt(1.0 - a * a)
    rp, rm = 1.0 + temp, 1.0 - temp
    eps = np.finfo(float).eps
    phks = ph + a / (2 * temp) * np.log(np.abs((r - rp + eps) / (r - rm + eps))) - np.arctan2(a, r)
    return r * np.sin(th) * np.cos(phks), r * np.sin(th) * np.sin(phks), r * np.cos(th)


def keplerian_zamo_velocity(a, r, th):
    """ZAMO spatial 4-velocity (r, phi, -theta axes) of the prograde circular orbit of angular
    velocity 1/(r^{3/2} + a) (Fit.jl:787-796)."""
    Om = 1.0 / (r * math.sqrt(r) + a)
    st, ct = math.sin(th), math.cos(th)
    S = r * r + a * a * ct * ct
    D = r * r - 2 * r + a * a
    Xi = (r * r + a * a) ** 2 - a * a * D * st * st
    gtt = -(1 - 2 * r / S)
    gtp = -2 * a * r * st * st / S
    gpp = Xi * st * st / S
    ut2 = -(gtt + 2 * gtp * Om + gpp * Om * Om)
    ut = 1.0 / math.sqrt(max(ut2, np.finfo(float).eps))
    omega = 2 * a * r / Xi
    sgpp = st * math.sqrt(Xi / S)
    return np.array([0.0, sgpp * (ut * Om - omega * ut), 0.0])


def _geometry(rng, a, n, single_frame=True):
    r = rng.uniform(4.0, 10.0, n)
    th = math.pi / 2 + rng.normal(0.0, 0.15, n)
    ph = rng.uniform(0.0, 2 * math.pi, n)
    x, y, z = quasi_cartesian_ks(a, r, th, ph)
    s = np.log(np.exp(rng.uniform(math.log(0.5), math.log(2.0), (3, n))))
    q = rng.normal(size=(4, n))
    q /= np.linalg.norm(q, axis=0)
    t0 = np.zeros(n)
    logw = np.full(n, math.log(1e9) if single_frame else math.log(40.0))
    omega = 1.0 / (r ** 1.5 + a)
    return r, th, np.vstack([x, y, z, s, q, t0, logw]), omega


def thin_splats(cfg, a, n):
    rng = seed(cfg)
    r, th, geo, omega = _geometry(rng, a, n)
    logA = np.log(np.exp(rng.uniform(math.log(0.1), math.log(1.0), n)))
    return np.vstack([geo, logA, omega])


def _fluid(rng, a, r, th, n):
    logB = rng.uniform(math.log(1.0), math.log(30.0), n)
    thB = np.arccos(rng.uniform(-1.0, 1.0, n))
    phB = rng.uniform(0.0, 2 * math.pi, n)
    frac = rng.uniform(0.7, 1.0, n)
    u = np.stack([keplerian_zamo_velocity(a, r[i], th[i]) * frac[i] for i in range(n)], axis=1)
    return logB, thB, phB, u


def thermal_splats(cfg, a, n, single_frame=True):
    rng = seed(cfg)
    r, th, geo, omega = _geometry(rng, a, n, single_frame)
    logne = rng.uniform(math.log(1e5), math.log(1e7), n)
    logTe = rng.uniform(math.log(3.0), math.log(50.0), n)
    logB, thB, phB, u = _fluid(rng, a, r, th, n)
    return np.vstack([geo, logne, logTe, logB, thB, phB, u, omega])


def powerlaw_splats(cfg, a, n, single_frame=True):
    rng = seed(cfg + 100)
    r, th, geo, omega = _geometry(rng, a, n, single_frame)
    logne = rng.uniform(math.log(1e3), math.log(1e5), n)
    p = rng.uniform(2.0, 3.5, n)
    lgmin = np.full(n, math.log(10.0))
    logB, thB, phB, u = _fluid(rng, a, r, th, n)
    return np.vstack([geo, logne, p, lgmin, logB, thB, phB, u, omega])


def kappa_splats(cfg, a, n, single_frame=True):
    rng = seed(cfg + 200)
    r, th, geo, omega = _geometry(rng, a, n, single_frame)
    logne = rng.uniform(math.log(1e4), math.log(1e6), n)
    kap = rng.uniform(3.0, 7.0, n)
    lw = rng.uniform(math.log(3.0), math.log(30.0), n)
    logB, thB, phB, u = _fluid(rng, a, r, th, n)
    return np.vstack([geo, logne, kap, lw, logB, thB, phB, u, omega])


def uv_disc(cfg, n, radius=10e9):
    rng = seed(cfg + 300)
    rr = radius * np.sqrt(rng.uniform(0.0, 1.0, n))
    ang = rng.uniform(0.0, 2 * math.pi, n)
    return rr * np.cos(ang), rr * np.sin(ang)


def length_unit():
    from .api import gravitational_radius
    return gravitational_radius(MASS_MSUN)


def pixel_size_rad(fov_lo, fov_hi, n):
    """Angular pixel size Δα L / D of Krang's grid (Δα = FOV/(n - 1))."""
    return (fov_hi - fov_lo) / (n - 1) * length_unit() / DISTANCE_CM


# ---- the benchmark configurations -------------------------------------------------------------------------
CONFIGS = {
    0: dict(name="cfg0-thin-64x64-N512", a=0.94, theta_o_deg=17.0, fov=12.0, res=64, N=512, ns=8, kind="thin",
            nus=[230e9], times=[0.0], nuv=0),
    1: dict(name="cfg1-polarized-128x128-N1024-32splats-512uv", a=0.5, theta_o_deg=60.0, fov=15.0, res=128, N=1024, ns=32,
            kind="thermal", nus=[230e9], times=[0.0], nuv=512),
}
