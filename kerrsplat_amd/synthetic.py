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


def boyer_lindquist(a, x, y, z):
    """BL (r, theta, phi) of a quasi-Cartesian Kerr-Schild point (the inverse of quasi_cartesian_ks,
    trajectories.jl:8-19)."""
    r = math.sqrt(x * x + y * y + z * z)
    th = math.acos(max(-1.0, min(1.0, z / r)))
    temp = math.sqrt(1.0 - a * a)
    rp, rm = 1.0 + temp, 1.0 - temp
    eps = np.finfo(float).eps
    ph = math.atan2(y, x) - a / (2 * temp) * math.log(abs((r - rp + eps) / (r - rm + eps))) + math.atan2(a, r)
    return r, th, ph


def coordinate_velocity(a, r, th, u):
    """(dr/dt, dtheta/dt, dphi/dt) of a fluid element with ZAMO spatial 4-velocity u (axes r, phi, -theta) at
    (r, theta) (trajectories.jl:21-26): the BL 4-velocity of the ZAMO components (gamma, u)."""
    st, ct = math.sin(th), math.cos(th)
    S = r * r + a * a * ct * ct
    D = r * r - 2 * r + a * a
    Xi = (r * r + a * a) ** 2 - a * a * D * st * st
    lapse, omega = math.sqrt(D * S / Xi), 2 * a * r / Xi
    sgrr, sgthth, sgpp = math.sqrt(S / D), math.sqrt(S), st * math.sqrt(Xi / S)
    g = math.sqrt(1.0 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    ut = g / lapse
    return np.array([u[0] / sgrr / ut, -u[2] / sgthth / ut, (u[1] / sgpp + omega * ut) / ut])


def trajectory_knots(a, params, times, substeps=4):
    """Knots 3 x nknots x ns of every splat's centre at the coordinate times (trajectory_knots,
    trajectories.jl:28-71): integrated from the centre at t0 (row 11) with the splat's ZAMO velocity (rows 18-20
    of a thermal column, 19-21 of a power-law or kappa column, 1-based) by RK4 in BL coordinates, substeps per
    knot interval, forward and backward from t0."""
    P = np.asarray(params)
    nr, ns = P.shape
    ru = 17 if nr == 21 else 18
    times = np.asarray(times, dtype=float)
    order = np.argsort(times)
    ts = times[order]
    dt_knot = abs(ts[-1] - ts[0]) / max(len(ts) - 1, 1)
    knots = np.zeros((3, len(times), ns))
    for i in range(ns):
        u = P[ru:ru + 3, i]
        t0 = P[10, i]
        q0 = np.array(boyer_lindquist(a, *P[0:3, i]))

        def f(q):
            return coordinate_velocity(a, q[0], q[1], u)

        def integrate(q, t_from, t_to):
            m = max(1, math.ceil(abs(t_to - t_from) / dt_knot * substeps)) if dt_knot > 0 else 1
            h = (t_to - t_from) / m
            for _ in range(m):
                k1 = f(q); k2 = f(q + h / 2 * k1); k3 = f(q + h / 2 * k2); k4 = f(q + h * k3)
                q = q + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            return q
        k0 = int(np.searchsorted(ts, t0))
        q, tp = q0, t0
        for k in range(k0, len(ts)):
            q = integrate(q, tp, ts[k]); tp = ts[k]
            knots[:, order[k], i] = quasi_cartesian_ks(a, q[0], q[1], q[2])
        q, tp = q0, t0
        for k in range(k0 - 1, -1, -1):
            q = integrate(q, tp, ts[k]); tp = ts[k]
            knots[:, order[k], i] = quasi_cartesian_ks(a, q[0], q[1], q[2])
    return knots


def stations(cfg, n, radius=5e9):
    """n station positions [wavelengths] for closure quantities: a disc of the given radius."""
    rng = seed(cfg + 400)
    rr = radius * np.sqrt(rng.uniform(0.0, 1.0, n))
    ang = rng.uniform(0.0, 2 * math.pi, n)
    return np.stack([rr * np.cos(ang), rr * np.sin(ang)], axis=1)


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
    # batched mixed precision: a 256^2 screen in uneven 32^2-128^2 tiles plus 32 screens of U[48, 160]^2, 3
    # frequencies x 16 frames over 100 M, 128 splats of the three populations (43 thermal, 43 power law, 42
    # kappa); N is not given by the configuration: 1024
    2: dict(name="cfg2-batched-256x256-uneven-tiles-32screens-N1024-128mixed-3nu-16frames", a=0.9, theta_o_deg=45.0, fov=15.0,
            res=256, N=1024, ns=128, kind="composite", nus=[86e9, 230e9, 345e9], times=list(np.linspace(0.0, 100.0, 16)), nuv=0,
            extra_screens=32, extra_size=(48, 160), tile_sizes=(32, 128)),
    # fwd + adjoint chi2 gradient: 64 thermal splats (1344 rows), 4 frames over 30 M, data rendered from the
    # parameters perturbed by N(0, 5%) plus noise at 1% of the peak, 2048 uv points, closures over 10 stations
    3: dict(name="cfg3-fwd-adjoint-chi2-128x128-N1024-64thermal-4frames-2048uv-10stations", a=0.7, theta_o_deg=30.0, fov=15.0,
            res=128, N=1024, ns=64, kind="thermal", nus=[230e9], times=[0.0, 10.0, 20.0, 30.0], nuv=2048, stations=10),
    # the full fit step: 1024 kappa splats whose centres follow 4 trajectory knots (from -100 to 200 M, covering the
    # 32 frames over 100 M and lookbacks of up to 100 M), 3 frequencies x 32 frames, 8192 uv points
    4: dict(name="cfg4-fit-step-512x512-N2048-1024kappa-4knots-3nu-32frames-8192uv", a=0.9, theta_o_deg=163.0, fov=15.0,
            res=512, N=2048, ns=1024, kind="kappa", nus=[86e9, 230e9, 345e9], times=list(np.linspace(0.0, 100.0, 32)), nuv=8192,
            knot_times=[-100.0, 0.0, 100.0, 200.0]),
}


def config_model(cfg):
    """The seeded splat parameters of a configuration: (kind, params) or, with knots, (kind, params, knots,
    knot_times); a composite configuration returns a list of (kind, params) parts."""
    c = CONFIGS[cfg]
    a, ns = c["a"], c["ns"]
    single = len(c["times"]) == 1
    if c["kind"] == "thin":
        return "thin", thin_splats(cfg, a, ns)
    if c["kind"] == "thermal":
        return "thermal", thermal_splats(cfg, a, ns, single_frame=single)
    if c["kind"] == "composite":
        n1 = (ns + 2) // 3
        n2 = (ns - n1 + 1) // 2
        return [("thermal", thermal_splats(cfg, a, n1, single_frame=single)), ("powerlaw", powerlaw_splats(cfg, a, n2, single_frame=single)),
                ("kappa", kappa_splats(cfg, a, ns - n1 - n2, single_frame=single))]
    P = kappa_splats(cfg, a, ns, single_frame=single)
    kt = np.asarray(c["knot_times"], dtype=float)
    return "kappa", P, trajectory_knots(a, P, kt), kt


def tile_splits(rng, res, tile_sizes):
    """Uneven tiles of a res x res screen: blocks of the largest tile size, each cut into four halves with
    probability 1/2 until the smallest size, as (a0, a1, b0, b1) half-open pixel ranges in screen order."""
    tmin, tmax = tile_sizes
    out = []

    def split(a0, b0, s):
        if s > tmin and rng.uniform() < 0.5:
            h = s // 2
            for db in (0, h):
                for da in (0, h):
                    split(a0 + da, b0 + db, h)
        else:
            out.append((a0, min(a0 + s, res), b0, min(b0 + s, res)))

    for b0 in range(0, res, tmax):
        for a0 in range(0, res, tmax):
            split(a0, b0, tmax)
    return out


def config_screens(cfg, res=None, extra=None, extra_size=None, tile_sizes=None):
    """The screens of a batched configuration: the main res x res screen over +-fov cut into uneven tiles
    (tile_splits) and `extra` square screens of sizes drawn from U[extra_size] over the same field of view.
    Returns (main camera, tile splits, extra cameras); the defaults are the configuration's sizes."""
    from .api import Camera
    c = CONFIGS[cfg]
    res = c["res"] if res is None else res
    extra = c["extra_screens"] if extra is None else extra
    lo, hi = c["extra_size"] if extra_size is None else extra_size
    ts = c["tile_sizes"] if tile_sizes is None else tile_sizes
    rng = np.random.Generator(np.random.PCG64(20261202 + 1000 * cfg + 7))
    fov = c["fov"]
    main = Camera((-fov, fov), (-fov, fov), res)
    splits = tile_splits(rng, res, ts)
    sizes = rng.integers(lo, hi + 1, size=extra)
    extras = [Camera((-fov, fov), (-fov, fov), int(n)) for n in sizes]
    return main, splits, extras
