"""Algorithmic work model of the benchmark configurations (test infrastructure).

Runs the oracle instantiated on a counting scalar type (oracle/ks_scalar.hpp `Counted`) over a
strided subset of each configuration's rays and writes the floating-point operations per
ray-sample to profiles/work_model.json: +, -, *, / and sqrt count 1, fma 2, every transcendental 1
(also tallied as "special"). bench.py reads that file to turn measured kernel time into an
achieved FLOP rate; nothing here is timed.

Run: python oracle/work_model.py
"""
import json
import math
import os
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


def main():
    out_path = os.path.join(ROOT, "profiles", "work_model.json")
    out = json.load(open(out_path)) if os.path.exists(out_path) else {}
    out[syn.CONFIGS[1]["name"]] = polarized_work(1)
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
