"""The oracle's transfer layer against fixtures and mathematics.

* coefficients against the symphony tables the reference commits (validation/symphony/*.csv,
  copied to tests/golden/symphony; the reference's gate is test/test_coefficients.jl:36-196 with
  tolerances 1e-12, and 1e-10 for kappa);
* K0, K1 and the kappa 2F1 factor against scipy;
* the 4x4 step against exp(-K D) and its integral by a 50-digit Taylor series with scaling and
  doubling (the reference's method, test/test_transfer_step.jl:22-39), on seeded random physical
  cases and the reference's thirteen special cases (test_transfer_step.jl:72-84);
* power-law rho_Q, rho_V against symphony's numerical rotativities and kappa j_V against symphony's numerical
  j_V, at the reference's looser gates for those fits;
* physical identities: Kirchhoff, V and rho_V odd under theta -> pi - theta, the polarization cap, the
  ZAMO redshift of a fluid at rest and the field-sign symmetry of the local frame. The frame of a moving
  fluid (boost, pitch angle, Walker-Penrose angle) is pinned end to end by ipole's RIAF images
  (tests/test_oracle_endtoend.py).
"""
import csv
import math
import os

import mpmath as mp
import numpy as np
import pytest
import scipy.special as sps

from conftest import GOLDEN

SYM = os.path.join(GOLDEN, "symphony")


def read_table(name):
    with open(os.path.join(SYM, name)) as f:
        rows = list(csv.DictReader(f))
    return {k.strip(): np.array([float(r[k]) for r in rows]) for k in rows[0]}


def relerr(a, b):
    return abs(a - b) / max(abs(b), 1e-280)


def test_thermal_vs_symphony(oracle):
    t = read_table("thermal_table.csv")
    names = ["jI", "jQ", "jV", "aI", "aQ", "aV", "rhoQ", "rhoV"]
    worst = dict.fromkeys(names + ["leung", "pjI", "pjQ", "pjV", "rhoV_dexter"], 0.0)
    for i in range(len(t["nu"])):
        args = (t["ne"][i], t["Thetae"][i], t["B"][i], t["nu"][i], t["theta"][i])
        c = oracle.thermal(*args)
        for k, n in enumerate(names):
            worst[n] = max(worst[n], relerr(c[k], t[n][i]))
        worst["leung"] = max(worst["leung"], relerr(oracle.thermal_leung(*args), t["jI_leung"][i]))
        pj = oracle.thermal_pandya(*args)
        worst["pjI"] = max(worst["pjI"], relerr(pj[0], t["jI_pandya"][i]))
        worst["pjQ"] = max(worst["pjQ"], relerr(pj[1], t["jQ_pandya"][i]))
        worst["pjV"] = max(worst["pjV"], relerr(pj[2], t["jV_pandya"][i]))
        cd = oracle.thermal(*args, dexter_rhoV=True)
        worst["rhoV_dexter"] = max(worst["rhoV_dexter"], relerr(cd[7], t["rhoV_dexter"][i]))
    for n, w in worst.items():
        assert w < 1e-12, (n, w)


def test_powerlaw_vs_symphony(oracle):
    t = read_table("powerlaw_table.csv")
    names = ["jI", "jQ", "jV", "aI", "aQ", "aV"]
    worst = dict.fromkeys(names, 0.0)
    for i in range(len(t["nu"])):
        c = oracle.powerlaw(t["ne"][i], t["p"][i], t["gamma_min"][i], t["gamma_max"][i], t["B"][i], t["nu"][i], t["theta"][i])
        for k, n in enumerate(names):
            worst[n] = max(worst[n], relerr(c[k], t[n][i]))
    for n, w in worst.items():
        assert w < 1e-12, (n, w)




def test_powerlaw_rotativities_vs_symphony(oracle):
    """Power-law rho_Q, rho_V (Jones & O'Dell via Dexter 2016) against symphony's numerical susceptibility
    integration (powerlaw_rho_table.csv) inside the closed forms' validity window nu > 3 nu_min, at the
    reference's gate (test_coefficients.jl:124-153): every sign agrees, rho_V within 5%, rho_Q within 35%."""
    t = read_table("powerlaw_rho_table.csv")
    ME, CL, EE = 9.1093826e-28, 2.99792458e10, 4.80320680e-10
    used = 0
    worstV = worstQ = 0.0
    for i in range(len(t["nu"])):
        nu, B, th, p, gmin, gmax, ne = (t[k][i] for k in ("nu", "B", "theta", "p", "gamma_min", "gamma_max", "ne"))
        if not (np.isfinite(t["rhoQ"][i]) and np.isfinite(t["rhoV"][i])):
            continue
        if not nu > 3 * gmin ** 2 * EE * B * math.sin(th) / (2 * math.pi * ME * CL):
            continue
        used += 1
        c = oracle.powerlaw(ne, p, gmin, gmax, B, nu, th)
        rQ, rV = c[6], c[7]
        assert np.sign(rQ) == np.sign(t["rhoQ"][i]) and np.sign(rV) == np.sign(t["rhoV"][i]), i
        worstV = max(worstV, abs(rV / t["rhoV"][i] - 1))
        worstQ = max(worstQ, abs(rQ / t["rhoQ"][i] - 1))
    assert used >= 10
    assert worstV < 0.05 and worstQ < 0.35, (worstV, worstQ)


def test_kappa_jv_vs_symphony_numerical(oracle):
    """kappa j_V (the fit) against symphony's numerical evaluation (kappa_jv_table.csv), the reference's gate
    (test_coefficients.jl:184-196): signs agree and the fit is within 50%."""
    t = read_table("kappa_jv_table.csv")
    worst = 0.0
    for i in range(len(t["nu"])):
        nu, kap, w, B, ne, th = (t[k][i] for k in ("nu", "kappa", "w", "B", "ne", "theta"))
        jV = oracle.kappa(ne, kap, w, B, nu, th, oracle.kappa_hyp(kap, w))[2]
        assert np.sign(jV) == np.sign(t["jV"][i]), i
        worst = max(worst, abs(jV / t["jV"][i] - 1))
    assert len(t["nu"]) >= 10 and worst < 0.5, worst


def test_kappa_vs_symphony(oracle):
    t = read_table("kappa_table.csv")
    names = ["jI", "jQ", "aI", "aQ", "aV", "rhoQ", "rhoV", "jV"]
    idx = {"jI": 0, "jQ": 1, "jV": 2, "aI": 3, "aQ": 4, "aV": 5, "rhoQ": 6, "rhoV": 7}
    worst = dict.fromkeys(names, 0.0)
    hyp = {}
    for i in range(len(t["nu"])):
        kap, w = t["kappa"][i], t["w"][i]
        if (kap, w) not in hyp:
            hyp[(kap, w)] = oracle.kappa_hyp(kap, w)
        c = oracle.kappa(t["ne"][i], kap, w, t["B"][i], t["nu"][i], t["theta"][i], hyp[(kap, w)])
        assert np.all(np.isfinite(c))
        for n in names:
            if n == "rhoQ" and kap == 5.0:
                continue          # both codes divide that interpolation by 5 instead of 3 (test_coefficients.jl:176)
            if n == "jV" and t["jV"][i] == 0:
                continue
            worst[n] = max(worst[n], relerr(c[idx[n]], t[n][i]))
    for n, w in worst.items():
        assert w < 1e-10, (n, w)


def test_kappa_hypergeometric_vs_scipy(oracle):
    for kap in (3.1, 3.5, 4.25, 5.0, 7.0):
        for w in (0.3, 3.0, 10.0, 30.0):
            ref = sps.hyp2f1(kap - 1 / 3, kap + 1, kap + 2 / 3, -kap * w)
            assert relerr(oracle.kappa_hyp(kap, w), ref) < 1e-12


def test_bessel_vs_scipy(oracle):
    for x in np.geomspace(0.003, 40, 60):
        k0, k1 = oracle.besselk01(x)
        assert relerr(k0, sps.k0(x)) < 1e-14 and relerr(k1, sps.k1(x)) < 1e-14


def test_thermal_identities(oracle):
    c = oracle.thermal(1e5, 10.0, 30.0, 230e9, 1.0)
    assert relerr(c[3], c[0] / oracle.planck(230e9, 10.0)) < 1e-14       # Kirchhoff
    assert c[1] > 0 and c[0] > c[1]
    cm = oracle.thermal(1e5, 10.0, 30.0, 230e9, math.pi - 1.0)
    assert c[2] > 0 and cm[2] < 0 and c[7] > 0 and cm[7] < 0
    c0 = oracle.thermal(1e5, 10.0, 30.0, 230e9, 0.0)
    assert c0[0] == 0 and c0[3] == 0 and c0[6] == 0 and c0[7] != 0
    cap = oracle.cap([1.0, 0.9, 0.9, 1.0, 0.1, 0.1, 0.0, 0.0])
    assert abs(math.hypot(cap[1], cap[2]) - 0.99) < 1e-15 and cap[4] == 0.1


# ---- the 4x4 step ------------------------------------------------------------------------------
def bigK(al, rh):
    aI, aQ, aU, aV = [mp.mpf(x) for x in al]
    rQ, rU, rV = [mp.mpf(x) for x in rh]
    return mp.matrix([[aI, aQ, aU, aV], [aQ, aI, rV, -rU], [aU, -rV, aI, rQ], [aV, rU, -rQ, aI]])


def big_step(K, D):
    """exp(-K D) and int_0^D exp(-K u) du: Taylor series after scaling D by 2^k, then doubling."""
    nrm = max(sum(abs(K[i, j]) for j in range(4)) for i in range(4))
    k = max(0, int(mp.ceil(mp.log(max(nrm * D, mp.mpf("1e-30")), 2))) + 4)
    d = D / mp.mpf(2) ** k
    O = mp.eye(4); term = mp.eye(4)
    E = d * mp.eye(4); eterm = d * mp.eye(4)
    for m in range(1, 60):
        term = term * (-K * d) / m
        O += term
        eterm = eterm * (-K * d) / (m + 1)
        E += eterm
    for _ in range(k):
        E = E + O * E
        O = O * O
    return O, E


def step_cases():
    rng = np.random.default_rng(20261202)
    cases = []
    for _ in range(120):
        j = rng.random(4) * np.array([1.0, 0.5, 0.5, 0.3])
        aI = 10.0 ** (rng.random() * 8 - 4)
        d = rng.standard_normal(3); d /= np.linalg.norm(d)
        fp = 0.99 * rng.random()
        al = np.concatenate([[aI], aI * fp * d])
        rd = rng.standard_normal(3); rd /= np.linalg.norm(rd)
        rh = 10.0 ** (rng.random() * 10 - 5) * rd
        D = 10.0 ** (rng.random() * 6 - 5)
        cases.append((j, al, rh, D))
    j = np.array([1.0, 0.3, -0.2, 0.1])
    special = [
        ([0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 0.7), ([2.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 0.7),
        ([0.0, 0.0, 0.0, 0.0], [3.0, -1.0, 2.0], 0.7), ([2.0, 1.2, -0.6, 0.9], [0.0, 0.0, 0.0], 0.7),
        ([2.0, 1.0, 0.0, 0.0], [0.0, 1.0, 0.0], 0.7), ([2.0, 1.0, 0.0, 0.0], [0.0, 1.0 + 1e-9, 0.0], 0.7),
        ([2.0, 1.0, 0.0, 0.0], [1e-6, 1.0, 0.0], 0.7), ([50.0, 40.0, 10.0, 5.0], [0.5, 0.2, 0.1], 30.0),
        ([1.0, 0.5, 0.1, 0.1], [1e4, 2e3, 5e3], 1.0), ([1e-9, 5e-10, 1e-10, 1e-10], [1e-8, 2e-9, 5e-9], 1e-3),
        ([1.0, 0.99, 0.0, 0.0], [0.0, 0.0, 0.0], 5.0), ([0.0, 0.0, 0.0, 0.0], [1e-216, 2e-218, 3e-162], 0.3),
        ([2.0, 0.0, 0.0, 0.0], [1e-170, 0.0, 1e-165], 0.3),
    ]
    for al, rh, D in special:
        cases.append((j, np.array(al), np.array(rh), D))
    return cases


@pytest.mark.slow
def test_step_vs_taylor(oracle):
    mp.mp.dps = 50
    worstO = worstE = 0.0
    for j, al, rh, D in step_cases():
        O, Ej = oracle.step_operator(al, rh, D)
        E = Ej @ j
        Ob, Eb = big_step(bigK(al, rh), mp.mpf(D))
        Obf = np.array([[float(Ob[i, k]) for k in range(4)] for i in range(4)])
        Ebf = np.array([[float(Eb[i, k]) for k in range(4)] for i in range(4)]) @ j
        nO = np.linalg.norm(Obf)
        eO = np.linalg.norm(O - Obf) / max(nO, 2.2e-308) if nO > 1e-300 else np.linalg.norm(O)
        eE = np.linalg.norm(E - Ebf) / max(np.linalg.norm(Ebf), 2.2e-308)
        scale = 1 + 0.02 * np.linalg.norm(rh) * D
        worstO = max(worstO, eO / scale); worstE = max(worstE, eE / scale)
    assert worstO < 1e-13 and worstE < 1e-13, (worstO, worstE)


def test_step_physics(oracle):
    # Kirchhoff equilibrium: S = j/alpha_I stays fixed for unpolarized coefficients
    al = np.array([2.0, 0.0, 0.0, 0.0]); rh = np.array([0.0, 0.0, 0.0])
    O, Ej = oracle.step_operator(al, rh, 0.7)
    j = np.array([3.0, 0.0, 0.0, 0.0])
    S = j / 2.0
    assert np.allclose(O @ S + Ej @ j, S, rtol=1e-14)
    # pure Faraday rotation rotates (Q, U) by angle rho_V D about V
    O, _ = oracle.step_operator(np.zeros(4), np.array([0.0, 0.0, 1.3]), 0.4)
    ang = 1.3 * 0.4
    S = O @ np.array([1.0, 1.0, 0.0, 0.0])
    assert abs(S[1] ** 2 + S[2] ** 2 - 1.0) < 1e-14 and abs(abs(math.atan2(S[2], S[1])) - ang) < 1e-14
    # semigroup: O(D1 + D2) = O(D1) O(D2)
    al = np.array([1.0, 0.3, -0.2, 0.4]); rh = np.array([2.0, -1.0, 0.5])
    O1, _ = oracle.step_operator(al, rh, 0.3)
    O2, _ = oracle.step_operator(al, rh, 0.5)
    O12, _ = oracle.step_operator(al, rh, 0.8)
    assert np.allclose(O1 @ O2, O12, rtol=1e-13, atol=1e-15)


def test_local_frame_properties(oracle):
    """The redshift of a static-in-ZAMO fluid equals the ZAMO redshift; chi is invariant under
    the field's sign; pitch angle cosine flips with B."""
    a, r, th = 0.9, 6.0, 1.1
    lam, eta = 2.0, 15.0
    u0 = np.zeros(3)
    B = np.array([0.3, 1.0, 0.4])
    g, cB, chi = oracle.local_frame(a, r, th, eta, lam, True, False, -2.2, 3.0, 0.8, u0, B)
    D = r * r - 2 * r + a * a; S = r * r + a * a * math.cos(th) ** 2
    Xi = (r * r + a * a) ** 2 - a * a * D * math.sin(th) ** 2
    gz = math.sqrt(D * S / Xi) / (1 - 2 * a * r / Xi * lam)
    assert relerr(g, gz) < 1e-14
    g2, cB2, chi2 = oracle.local_frame(a, r, th, eta, lam, True, False, -2.2, 3.0, 0.8, u0, -B)
    assert abs(cB2 + cB) < 1e-14 and abs(math.remainder(chi2 - chi, math.pi)) < 1e-12
