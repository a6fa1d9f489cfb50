"""Golden vectors for the geodesic layer, from an independent method.

Nothing here uses a closed form: every value comes from mpmath at 40 digits by
  * the real-root count from mpmath.polyroots of the radial quartic,
  * the total Mino time from mpmath.quad of 1/sqrt(R(r)) (to the turning point and back for a
    ray that turns outside the horizon, to the outer horizon otherwise),
  * r, theta, t~ and phi at the sample times tau_k = k tau_total/(N + 1) by integrating the
    geodesic equations in Mino time with mpmath.odefun (Taylor series):
        u = 1/r:   u'' = Q'(u)/2,  Q(u) = u^4 R(1/u) = 1 + A u^2 + B u^3 + C u^4,  u(0) = 0, u'(0) = 1
        theta:     theta'' = Theta'(theta)/2,  theta(0) = theta_o,  theta'(0) = -sign(beta) sqrt(Theta(theta_o))
        w = t~ + r + 2 ln r:  w' = dt/dtau - (1 + 2u) u'/u^2 (finite at u = 0),  w(0) = 0
        phi:       phi' = a(2r - a lambda)/Delta + lambda/sin^2 theta,  phi(0) = 0
    which is smooth through the radial and polar turning points. The observer-side start
    u = 0 is taken at tau_0 = 1e-16 from its Taylor series.
This restates only the physics (the Kerr null geodesic equations, G&L 2020 eqs. 7-9), the
regularisation of t~ (r_o + 2 ln r_o subtracted as r_o -> infinity) and the sign convention
of beta; the reference's own gate 3 (test/highprec_reference.jl:144-221) uses quadrature of
the same Mino-time rates.

Run: python tests/golden/make_geodesic_golden.py  (a few minutes; writes geodesic_golden.json)
"""
import json
import os
import sys
from multiprocessing import Pool

import mpmath as mp

DPS = 40
N = 8

# (a, theta_o in degrees, alpha, beta): the configuration spacetimes and the reference's gate
# spacetimes, screen points chosen to cover four, two and no real roots, vortical rays,
# observers beyond 90 degrees and negative beta
PIXELS = []
for a, th in [(0.94, 17.0), (0.5, 60.0), (0.9, 45.0), (0.7, 30.0), (0.9, 163.0), (0.3, 75.0), (0.99, 80.0)]:
    for al, be in [(3.0, 5.0), (-6.0, 1.0), (0.2, 0.1), (-1.0, 0.5), (0.5, 1.5), (-2.0, -3.0), (4.0, -3.0),
                   (9.0, -7.0), (-0.7, -0.3)]:
        PIXELS.append((a, th, al, be))


def golden_pixel(args):
    a, thdeg, al, be = args
    mp.mp.dps = DPS
    a = mp.mpf(a); tho = mp.radians(mp.mpf(thdeg)); al = mp.mpf(al); be = mp.mpf(be)
    lam = -al * mp.sin(tho)
    eta = (al ** 2 - a ** 2) * mp.cos(tho) ** 2 + be ** 2
    A = a * a - eta - lam * lam
    B = 2 * (eta + (lam - a) ** 2)
    C = -a * a * eta
    rts = mp.polyroots([1, 0, A, B, C], maxsteps=200, extraprec=200)
    scale = max(1, max(abs(r) for r in rts))
    real = sorted([mp.re(r) for r in rts if abs(mp.im(r)) < mp.mpf(10) ** -25 * scale])
    nreal = len(real)
    rp = 1 + mp.sqrt(1 - a * a)
    Rr = lambda r: r ** 4 + A * r * r + B * r + C
    if nreal == 4 and real[3] > rp:
        r4 = real[3]
        fo = mp.re(mp.quad(lambda r: 1 / mp.sqrt(Rr(r)), [r4, r4 + 1, 2 * r4 + 10, mp.inf]))
        tau_total = 2 * fo
    else:
        tau_total = mp.re(mp.quad(lambda r: 1 / mp.sqrt(Rr(r)), [rp, rp + 1, 2 * rp + 10, mp.inf]))
        fo = None

    Th = lambda th: eta + a * a * mp.cos(th) ** 2 - lam * lam * mp.cot(th) ** 2
    dTh = lambda th: -2 * a * a * mp.cos(th) * mp.sin(th) + 2 * lam * lam * mp.cos(th) / mp.sin(th) ** 3
    sb = 1 if be >= 0 else -1
    th0p = -sb * mp.sqrt(max(Th(tho), 0))

    def rhs(tau, y):
        u, up, th, thp, w, ph = y
        du = up
        dup = A * u + mp.mpf(3) / 2 * B * u * u + 2 * C * u ** 3
        dth = thp
        dthp = dTh(th) / 2
        s2 = mp.sin(th) ** 2
        Dl_u2 = 1 - 2 * u + a * a * u * u              # Delta u^2
        # dt/dtau = (r^2+a^2)(r^2+a^2-a lam)/Delta + a(lam - a s2), written in u
        rate_t = (1 + a * a * u * u) * (1 + (a * a - a * lam) * u * u) / (u * u * Dl_u2) + a * (lam - a * s2)
        dw = rate_t - (1 + 2 * u) * up / (u * u)
        dph = a * u * (2 - a * lam * u) / Dl_u2 + lam / s2
        return [du, dup, dth, dthp, dw, dph]

    t0 = mp.mpf(10) ** -16
    y0 = [t0, mp.mpf(1), tho + th0p * t0, th0p + dTh(tho) / 2 * t0, mp.mpf(0), mp.mpf(0)]
    # first-order start for w and phi at t0 (their rates are finite at u = 0)
    r0 = rhs(t0, y0)
    y0[4] = r0[4] * t0
    y0[5] = r0[5] * t0
    sol = mp.odefun(rhs, t0, y0, tol=mp.mpf(10) ** -(DPS - 8), degree=30)
    samples = []
    for k in range(1, N + 1):
        tau = tau_total * k / (N + 1)
        u, up, th, thp, w, ph = sol(tau)
        r = 1 / u
        t = w - r - 2 * mp.log(r)
        samples.append({"tau": float(tau), "t": float(t), "r": float(r), "theta": float(th), "phi": float(ph),
                        "nur": bool(up > 0), "nuth": bool(thp < 0)})
    out = {"a": float(a), "theta_o_deg": thdeg, "alpha": float(al), "beta": float(be),
           "lam": float(lam), "eta": float(eta), "nreal": nreal, "tau_total": float(tau_total),
           "I0_inf": float(fo) if fo is not None else None,
           "roots": [[float(mp.re(r)), float(mp.im(r))] for r in rts], "N": N, "samples": samples}
    print(f"done a={float(a)} th={thdeg} al={float(al)} be={float(be)} nreal={nreal}", flush=True)
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    nproc = int(os.environ.get("GOLDEN_PROCS", "6"))
    with Pool(nproc) as pool:
        res = pool.map(golden_pixel, PIXELS)
    with open(os.path.join(here, "geodesic_golden.json"), "w") as f:
        json.dump({"method": "mpmath odefun on the Mino-time geodesic equations, 40 digits", "pixels": res}, f, indent=1)


if __name__ == "__main__":
    main()
