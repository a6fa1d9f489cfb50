"""Algorithmic work model of the benchmark configurations (test infrastructure).

Counts the floating-point operations per ray-sample of each pipeline stage and writes them to
profiles/work_model.json: +, -, *, / and sqrt count 1, fma 2, every transcendental 1 (also tallied as
"special"). bench.py reads that file to turn measured kernel time into an achieved FLOP rate;
nothing here is timed.

* The thermal fit step (128 x 128, 1024 samples, 32 thermal splats): the oracle instantiated on its
  counting scalar (oracle/ks_scalar.hpp `Counted`) over a strided subset of the rays, phase by phase.
* The full fit step (512 x 512, 2048 samples, 1024 kappa splats on trajectory knots): the march by the
  oracle's counting scalar; the transfer and the adjoint by the device kernels' own per-sample and
  per-pair templates on a counting scalar (oracle/ks_workmodel.hip), fed with the kept (sample, splat)
  pairs of a random subset of rays at one (frame, frequency) slice (weight above the cutoff 1e-6, the
  sample outside the horizon guard), found here with numpy from the oracle's samples. The counted work
  is that of the kept pairs and of their samples; the per-tile and per-unit cull tests are not counted.

Run: python oracle/work_model.py [thermal|kappa]
"""
import json
import math
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from kerrsplat_amd import synthetic as syn  # noqa: E402
from kerrsplat_amd.api import Camera  # noqa: E402


def polarized_work(cfg_id, stride=16):
    """Per ray-sample work of the fwd+adjoint step, by phase: the march (t~, r, theta, phi), the
    forward transfer (cull + splat sums + step operators), the tails, and the adjoint sweep split
    into its per-sample step adjoints and its per-splat element adjoints."""
    cfg = syn.CONFIGS[cfg_id]
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    cam = Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    idx = np.arange(0, cam.npix, stride)
    P = syn.thermal_splats(cfg_id, a, cfg["ns"])
    ds = np.random.default_rng(20261202).normal(size=(len(idx), 4))
    c, sp = oracle.count_gradient(a, th, cam.alphas[idx], cam.betas[idx], N, P, 0.0, cfg["nus"][0], syn.length_unit(), ds)
    n = len(idx) * N
    phases = {"march": (c[0], sp[0]), "transfer": (c[1], sp[1]), "tails": (c[2], sp[2]),
              "adjoint_steps": (c[4], sp[4]), "adjoint_elements": (c[3] - c[4], sp[3] - sp[4])}
    out = {k: {"flops_per_ray_sample": v[0] / n, "special_per_ray_sample": v[1] / n} for k, v in phases.items()}
    fwd = (c[0] + c[1]) / n
    return {"flops_per_ray_sample": fwd, "flops_per_ray_sample_fwd_adjoint": c[:4].sum() / n, "phases": out,
            "rays_counted": int(len(idx)), "samples_per_ray": N, "splats": cfg["ns"],
            "method": "oracle dual sweep on the Counted scalar type, every %d-th ray of the screen" % stride,
            "hbm_bytes_per_launch": None}


def kept_pairs(cfg, P, kn, kt, al, be, N, t_obs):
    """The oracle's samples of the rays (alpha, beta) and, per valid sample (flag bit 1, outside the horizon
    guard), the kept splats: Gaussian weight above the cutoff at the knot-interpolated centre (linear between the
    knots, held at the ends), the transfer's test. Returns (pix4 per sample, smp7 per sample, kept lists) over the
    valid samples in ray order."""
    import oracle
    a, th = cfg["a"], math.radians(cfg["theta_o_deg"])
    smp = oracle.samples(a, th, al, be, N)
    r, thk, ph, fl, t = smp["r"], smp["theta"], smp["phi"], smp["flags"], smp["t"]
    hor = (1.0 + math.sqrt(1.0 - a * a)) * (1.0 + 1e-3)
    ok = ((fl & 1) > 0) & (r > hor)
    x, y, z = syn.quasi_cartesian_ks(a, r, thk, ph)
    te = t_obs - t
    ns = P.shape[1]
    keep = np.zeros((ns,) + r.shape, dtype=bool)
    for i in range(ns):
        tt = np.clip(te, kt[0], kt[-1])
        k = np.clip(np.searchsorted(kt, tt, side="right") - 1, 0, len(kt) - 2)
        w = (tt - kt[k]) / (kt[k + 1] - kt[k])
        c = [(1 - w) * kn[d, k, i] + w * kn[d, k + 1, i] for d in range(3)]
        q = P[6:10, i] / np.linalg.norm(P[6:10, i])
        qw, qx, qy, qz = q
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)],
                      [2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)],
                      [2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)]])
        d = np.stack([x - c[0], y - c[1], z - c[2]], axis=-1)
        u = (d @ R) * np.exp(-P[3:6, i])
        g = (u ** 2).sum(-1) + ((te - P[10, i]) * math.exp(-P[11, i])) ** 2
        keep[i] = ok & (np.exp(-0.5 * g) > 1e-6)
    sin_o, cos_o = math.sin(th), math.cos(th)
    eta = (al ** 2 - a * a) * cos_o ** 2 + be ** 2
    lam = -al * sin_o
    pix = np.stack([al, be, eta, lam], axis=1)
    rows, cols = np.nonzero(ok)
    pix4 = pix[rows]
    smp7 = np.stack([t[rows, cols], r[rows, cols], thk[rows, cols], x[rows, cols], y[rows, cols], z[rows, cols],
                     fl[rows, cols].astype(float)], axis=1)
    lists = [np.nonzero(keep[:, i, k])[0] for i, k in zip(rows, cols)]
    return pix4, smp7, lists, int(ok.sum()), int(r.size)


def kappa_knot_work(cfg_id=4, nrays=96, count_samples=1500, seed=11):
    """Per ray-sample work of the full fit step's slice (kappa splats on knots): march, transfer (splat sums over
    the kept splats, the FP64 sample geometry, the step), the adjoint's per-sample steps and its element adjoints
    (each kept pair: the weight's log-derivatives and the element adjoint over the sample geometry the transfer
    stored), split into FP64 and scalar-of-the-path (FP32 on the mixed path) operations."""
    import ctypes
    import oracle
    cfg = syn.CONFIGS[cfg_id]
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    _, P, kn, kt = syn.config_model(cfg_id)
    t_obs, nu, L = cfg["times"][len(cfg["times"]) // 2], cfg["nus"][1], syn.length_unit()
    cam = Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    rng = np.random.default_rng(seed)
    idx = rng.choice(cam.npix, nrays, replace=False)
    al, be = cam.alphas[idx], cam.betas[idx]
    pix4, smp7, lists, nvalid, nall = kept_pairs(cfg, P, kn, kt, al, be, N, t_obs)
    counts = np.array([len(l) for l in lists])
    active = counts > 0
    pairs = int(counts.sum())
    # the march, by the oracle's counting scalar
    fm, sm = ctypes.c_double(), ctypes.c_double()
    oracle.lib().kso_count_march(ctypes.c_double(a), ctypes.c_double(th), ctypes.c_int(nrays), oracle._p(oracle._f64(al)),
                                 oracle._p(oracle._f64(be)), ctypes.c_int(N), ctypes.c_int(64), ctypes.byref(fm), ctypes.byref(sm))
    # the device templates over a random subset of the active samples
    act = np.nonzero(active)[0]
    pick = np.sort(rng.choice(act, min(count_samples, len(act)), replace=False))
    koff = np.zeros(len(pick) + 1, dtype=np.int32)
    koff[1:] = np.cumsum([len(lists[s]) for s in pick])
    kept = np.concatenate([lists[s] for s in pick]).astype(np.int32)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "workmodel"], check=True)
    wm = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "libks_workmodel.so"))
    Pf = np.asfortranarray(P)
    knf = np.asfortranarray(kn)
    ktf = np.ascontiguousarray(kt, dtype=np.float64)
    p4 = np.ascontiguousarray(pix4[pick])
    s7 = np.ascontiguousarray(smp7[pick])
    res_ = {}
    for label, fp32 in (("mixed", 1), ("fp64", 0)):
        out = np.zeros(12)
        rc = wm.kswm_count_slice(ctypes.c_int(3), ctypes.c_int(22), ctypes.c_int(P.shape[1]), Pf.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_double(1e5), knf.ctypes.data_as(ctypes.c_void_p), ktf.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_int(len(kt)), ctypes.c_double(a), ctypes.c_double(th), ctypes.c_double(t_obs),
                                 ctypes.c_double(nu), ctypes.c_double(L), ctypes.c_int(len(pick)), p4.ctypes.data_as(ctypes.c_void_p),
                                 s7.ctypes.data_as(ctypes.c_void_p), koff.ctypes.data_as(ctypes.c_void_p),
                                 kept.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(fp32), out.ctypes.data_as(ctypes.c_void_p))
        if rc:
            raise RuntimeError(f"kswm_count_slice failed ({rc})")
        res_[label] = out
    n_cnt, p_cnt = len(pick), int(koff[-1])
    rs = nall                                # ray-samples of the counted rays
    f_active = active.sum() / rs             # active samples per ray-sample
    f_pairs = pairs / rs                     # kept pairs per ray-sample

    def phases(out):
        per_s = lambda k: out[k] / n_cnt
        per_p = lambda k: out[k] / p_cnt
        tr32 = f_active * (per_s(0) + per_s(4))
        tr64 = f_active * per_s(2)
        st32 = f_active * per_s(6)
        el32 = f_pairs * per_p(8)
        el64 = f_pairs * per_p(10)
        sp = lambda k, f: f * out[k] / (n_cnt if k < 8 else p_cnt)
        return {
            "march": {"flops_per_ray_sample": fm.value / rs, "fp64_flops_per_ray_sample": fm.value / rs, "special_per_ray_sample": sm.value / rs},
            "transfer": {"flops_per_ray_sample": tr32 + tr64, "fp64_flops_per_ray_sample": tr64,
                         "special_per_ray_sample": sp(1, f_active) + sp(3, f_active) + sp(5, f_active)},
            "adjoint_steps": {"flops_per_ray_sample": st32, "fp64_flops_per_ray_sample": 0.0, "special_per_ray_sample": sp(7, f_active)},
            "adjoint_elements": {"flops_per_ray_sample": el32 + el64, "fp64_flops_per_ray_sample": el64,
                                 "special_per_ray_sample": sp(9, f_pairs) + sp(11, f_pairs)},
        }
    return {"phases": phases(res_["mixed"]), "phases_fp64_path": phases(res_["fp64"]),
            "flops_per_pair": {"element_adjoint": res_["mixed"][8] / p_cnt, "geometry_fp64_per_sample": res_["mixed"][2] / n_cnt,
                               "transfer_splat_sum": res_["mixed"][0] / p_cnt},
            "pairs_per_ray_sample": f_pairs, "active_samples_per_ray_sample": f_active, "valid_samples_per_ray_sample": nvalid / rs,
            "rays_counted": int(nrays), "samples_counted": n_cnt, "pairs_counted": p_cnt, "samples_per_ray": N, "splats": int(P.shape[1]),
            "slice": {"t_obs": t_obs, "nu": nu},
            "method": ("march: the oracle's counting scalar over the rays; transfer and adjoint: the device kernels' templates "
                       "on a counting scalar (oracle/ks_workmodel.hip) over a random subset of the active samples of %d random "
                       "rays and their kept pairs; fp64_flops are the FP64 share on the mixed path (the sample geometry)" % nrays),
            "hbm_bytes_per_launch": None}


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    out_path = os.path.join(ROOT, "profiles", "work_model.json")
    out = json.load(open(out_path)) if os.path.exists(out_path) else {}
    if which in ("all", "thermal"):
        out[syn.CONFIGS[1]["name"]] = polarized_work(1)
    if which in ("all", "kappa"):
        out[syn.CONFIGS[4]["name"]] = kappa_knot_work(4)
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
