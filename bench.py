"""Benchmark: ray-samples/s of the geodesic march + polarized transfer + dual-sweep adjoint on one MI355X.

Workload (the configuration the headline metric is quoted on that fits one GPU): a = 0.5,
theta_o = 60 deg, 128 x 128 screen over +-15 M, 1024 Mino-time samples per ray, 32 thermal splats,
full IQUV at 230 GHz, visibilities on 512 uv points (seeded synthetic draws). The per-pixel
constants (K0/K1) are built once beforehand and the splat parameters are resident on the device.
One fwd+adjoint step is the device pipeline of a visibility fit iteration (frame_chi2_seed! between
the two passes of the dual sweep): the Recurrence(64) march of every ray into sample storage, the
polarized transfer with the adjoint records, the visibilities of the image with their χ² against
fixed synthetic data and the adjoint seed of the weighted residuals (left on the device), and the
reverse pass (the gradient with respect to all 21 x 32 splat parameters). The forward-only step
(march + transfer + visibilities) is timed beside it.

Prints one JSON line; value = ray-samples per second (pixels x samples x frames x frequencies / s)
of the fwd+adjoint step.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ray-samples/s fwd+adjoint, mixed FP32 vs plain FP64 on 1 GPU @ ≤1e-4 Stokes err"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    p.add_argument("--copy-probe", action="store_true",
                   help="also run device-to-device copies of 256 MiB (the reference for scaling HBM byte counters)")
    return p.parse_args()


def copy_probe(nbytes=1 << 28, reps=3):
    """Device-to-device copies of nbytes through the HIP runtime (its blit kernel): the kernel of known
    traffic (nbytes read + nbytes written per copy) that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    hip.hipMemset(a, 1, n)
    for _ in range(reps):
        hip.hipMemcpy(b, a, n, 3)                      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(cfg, P, L, target_s, ds):
    """The oracle (FP64 C++ restatement of the reference's per-ray CPU path: Recurrence(64) march,
    transfer and dual sweep, OpenMP over all host cores) on a bounded random subset of the same rays
    with the same seeds; fwd+adjoint ray-samples/s."""
    import oracle
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    from kerrsplat_amd import Camera
    cam = Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    dso = ds[:, :, 0, 0, :].reshape(cam.npix, 4, order="F")
    threads = oracle.num_threads()
    rng = np.random.default_rng(7)
    order = rng.permutation(cam.npix)
    n = 64
    t0 = time.perf_counter()
    oracle.polarized_gradient(a, th, cam.alphas[order[:n]], cam.betas[order[:n]], N, P, 0.0, cfg["nus"][0], L, dso[order[:n]])
    dt = time.perf_counter() - t0
    n2 = int(min(cam.npix, max(n, n * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    oracle.polarized_gradient(a, th, cam.alphas[order[:n2]], cam.betas[order[:n2]], N, P, 0.0, cfg["nus"][0], L, dso[order[:n2]])
    dt2 = time.perf_counter() - t0
    return {"value": n2 * N / dt2, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"{n2} of {cam.npix} rays (random subset) x {N} samples, fwd+adjoint (dual sweep), FP64 oracle, {dt2:.1f} s"}


def visibility_stage(K, syn, cfg, model, L, res=512, nuv=8192, reps=3):
    """The visibility transform's matrix-core product at the largest configuration's size (512 x 512
    screen, 8192 uv points, one frame; the image rendered from the same splats over a fused 64-sample
    march): HIP-event duration of the FP64 GEMM and its rate against the FP64 matrix peak."""
    a, th = cfg["a"], math.radians(cfg["theta_o_deg"])
    cam = K.Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    cache = K.GeodesicCache(cam, 64, store_samples=False)
    K.regenerate(cache, a, th, marcher=K.Fused(64))
    K.render_device(cache, model, cfg["times"][:1], cfg["nus"][:1], L)
    u, v = syn.uv_disc(4, nuv)
    psize = syn.pixel_size_rad(-cfg["fov"], cfg["fov"], res)
    K.visibilities(cache, 1, psize, u, v)
    ms = []
    for _ in range(reps):
        K.visibilities(cache, 1, psize, u, v)
        ms.append(K.stage_ms("vis"))
    t = float(np.median(ms))
    flops = 2.0 * (2 * nuv) * res * (4 * res)          # C2 = [cos; sin](2 nuv x n_alpha) . image (n_alpha x 4 n_beta)
    achieved = flops / (t * 1e-3) / 1e12
    return {"workload": f"{res}x{res} screen, {nuv} uv, 1 frame", "kernel": "ks_dgemm<false>", "bound": "mfma-fp64",
            "ms": t, "achieved_tflops": achieved, "peak_tflops": FP64_PEAK_TFLOPS, "frac": achieved / FP64_PEAK_TFLOPS}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus N > 1 is not supported "
                                   "(the sharding design is described in DESIGN.md)"}))
        sys.exit(2)
    import kerrsplat_amd as K
    from kerrsplat_amd import api, synthetic as syn

    cfg = syn.CONFIGS[args.config]
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    cam = K.Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    L = syn.length_unit()
    P = syn.thermal_splats(args.config, a, cfg["ns"])
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, a, th, marcher=K.Recurrence(64))
    model = K.Model(P, kind="thermal")
    ds = np.random.default_rng(20261202).normal(size=(res, res, len(cfg["times"]), len(cfg["nus"]), 4))
    u, v = syn.uv_disc(args.config, cfg["nuv"])
    psize = syn.pixel_size_rad(-cfg["fov"], cfg["fov"], res)
    lib, ctx = api._load(), api.context()
    times, nus = cfg["times"], cfg["nus"]
    nframes = len(times) * len(nus)
    # fixed synthetic visibility data: the FP64 model's visibilities, perturbed by 5 % and thermal noise
    rng = np.random.default_rng(20261203)
    K.precision(cache, "fp64")
    K.render_device(cache, model, times, nus, L)
    V0 = K.visibilities(cache, nframes, psize, u, v) if cfg["nuv"] else None
    if cfg["nuv"]:
        sig = 0.01 * np.abs(V0[..., 0]).max() * np.ones(V0.shape)
        vdata = V0 * (1 + 0.05 * rng.normal(size=V0.shape)) + sig * (rng.normal(size=V0.shape) + 1j * rng.normal(size=V0.shape))
    dparams = np.zeros(P.shape, order="F")
    chis = []

    def forward_step():
        K.march(cache)
        K.render_device(cache, model, times, nus, L)
        if cfg["nuv"]:
            K.visibilities(cache, nframes, psize, u, v)

    def adjoint_step():
        K.march(cache)
        if not cfg["nuv"]:
            K.polarized_gradient(cache, model, times, nus, L, ds, dparams=dparams)
            return
        K.gradient_forward(cache, model, times, nus, L, copy=False)
        chis.append(sum(K.vis_chi2_seed(cache, f, psize, u, v, vdata[f], sig[f]) for f in range(nframes)))
        K.gradient_reverse(cache, model, dparams)

    def timed(step, stages):
        """W untimed warmups, then K steps bracketed by stream synchronizations; per step, each listed
        stage's kernel duration from the HIP events the library records around its launch."""
        for _ in range(args.warmup):
            step()
        lib.ks_ctx_sync(ctx)
        ms = {k: [] for k in stages}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            for k in stages:
                ms[k].append(api.stage_ms(k))
        lib.ks_ctx_sync(ctx)
        return (time.perf_counter() - t0) / args.steps, {k: float(np.mean(v)) for k, v in ms.items()}

    STG_FWD, STG_ADJ = ["march", "transfer"], ["march", "transfer", "adjoint_steps", "adjoint_elements"]
    work = cam.npix * N * len(times) * len(nus)
    runs = {}
    for prec in ("fp64", "mixed"):
        K.precision(cache, prec)
        fwd_s, fwd_ms = timed(forward_step, STG_FWD)
        adj_s, adj_ms = timed(adjoint_step, STG_ADJ)
        dparams[:] = 0.0
        adjoint_step()
        cube = K.gradient_forward(cache, model, times, nus, L)
        runs[prec] = {"fwd_s": fwd_s, "fwd_ms": fwd_ms, "adj_s": adj_s, "adj_ms": adj_ms, "grad": dparams.copy(), "cube": cube,
                      "chi2": chis[-1] if chis else None}
    img_err = float(np.abs(runs["mixed"]["cube"] - runs["fp64"]["cube"]).max() / np.abs(runs["fp64"]["cube"][..., 0]).max())
    grad_err = float(np.abs(runs["mixed"]["grad"] - runs["fp64"]["grad"]).max() / np.abs(runs["fp64"]["grad"]).max())
    head = "mixed" if img_err <= 1e-4 else "fp64"        # the mixed path counts only within its error bar
    adj_s, adj_ms = runs[head]["adj_s"], runs[head]["adj_ms"]
    ms_per_step = 1e3 * adj_s
    value = work / adj_s

    # per-stage rooflines from the algorithmic work model (oracle on the Counted scalar): VALU bound,
    # against the FP64 vector peak on the FP64 path and the FP32 vector peak on the mixed path
    wm_path = os.path.join(ROOT, "profiles", "work_model.json")
    wm = json.load(open(wm_path)).get(cfg["name"]) if os.path.exists(wm_path) else None
    tpl = {"fp64": "double", "mixed": "f32"}
    cull = "true" if 64 <= cfg["ns"] <= 1024 else "false"      # the transfer's per-tile cull (from 64 to 1024 splats)
    kernel_of = {"march": lambda p_: "ks_march_store<1>", "transfer": lambda p_: f"ks_chunk_transfer<1, {tpl[p_]}, {cull}>",
                 "adjoint_steps": lambda p_: f"ks_adj_segment<{tpl[p_]}>", "adjoint_elements": lambda p_: f"ks_adj_splats<{tpl[p_]}>"}
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE summary (rocprofv3 --pmc, scaled by a copy probe)
    tr_path = os.path.join(ROOT, "profiles", "traffic_128x128_N1024_32thermal.json")
    traffic = json.load(open(tr_path))["kernels"] if (os.path.exists(tr_path) and args.config == 1) else {}

    def stage_table(prec):
        peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        out = {}
        for k, t_ms in runs[prec]["adj_ms"].items():
            st = {"kernel": kernel_of[k](prec), "ms": t_ms}
            if st["kernel"] in traffic:
                st["hbm_bytes_per_launch"] = traffic[st["kernel"]]["bytes"]
            if wm and "phases" in wm:
                ph = wm["phases"]
                per = {"march": ph["march"]["flops_per_ray_sample"], "transfer": ph["transfer"]["flops_per_ray_sample"],
                       "adjoint_steps": ph["adjoint_steps"]["flops_per_ray_sample"] + ph["tails"]["flops_per_ray_sample"],
                       "adjoint_elements": ph["adjoint_elements"]["flops_per_ray_sample"]}[k]
                units = cam.npix * N if k == "march" else work
                achieved = per * units / (t_ms * 1e-3) / 1e12
                pk = FP64_PEAK_TFLOPS if k == "march" else peak     # the geodesics are FP64 on both paths
                st.update({"flops_per_ray_sample": per, "achieved_tflops": achieved, "peak_tflops": pk, "frac": achieved / pk})
            out[k] = st
        return out

    stages = {p_: stage_table(p_) for p_ in runs}
    roofline = None
    dom = max(adj_ms, key=adj_ms.get)
    d = stages[head][dom]
    if "achieved_tflops" in d:
        roofline = {"bound": "valu-fp64" if (head == "fp64" or dom == "march") else "valu-fp32", "achieved": d["achieved_tflops"],
                    "peak": d["peak_tflops"], "unit": "TFLOP/s", "frac": d["frac"], "traffic": d.get("hbm_bytes_per_launch"), "kernel": d["kernel"],
                    "kernel_ms": d["ms"], "flops_per_ray_sample": d["flops_per_ray_sample"]}
    paths = {p_: {"fwd_adjoint": {"value": work / r["adj_s"], "unit": "ray-samples/s", "ms_per_step": 1e3 * r["adj_s"],
                                  "gradient_evals_per_s": 1.0 / r["adj_s"], "stage_ms": r["adj_ms"]},
                  "forward": {"value": work / r["fwd_s"], "unit": "ray-samples/s", "ms_per_step": 1e3 * r["fwd_s"], "stage_ms": r["fwd_ms"]}}
             for p_, r in runs.items()}

    vis_stage = visibility_stage(K, syn, cfg, model, L) if cfg["nuv"] else None
    if args.copy_probe:
        copy_probe()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, P, L, args.cpu_seconds, ds)
    out = {"metric": METRIC, "value": value, "unit": "ray-samples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "mixed f32/f64" if head == "mixed" else "f64",
           "data": "synthetic (seeded splat draws and uv disc)",
           "config": {"workload": cfg["name"], "a": a, "theta_o_deg": cfg["theta_o_deg"], "screen": [res, res],
                      "samples_per_ray": N, "splats": cfg["ns"], "kind": cfg["kind"], "frames": len(times),
                      "frequencies": len(nus), "uv_points": cfg["nuv"],
                      "path": ("mixed: FP64 geodesics (Recurrence(64) into stored samples, anchor segments in parallel), "
                               "FP32 transfer and dual-sweep adjoint in 32-sample chunks (FP64 step above Faraday depth 10), "
                               "visibility chi2 and seed on the device (FP64 GEMMs)")
                              if head == "mixed" else
                              "plain FP64: Recurrence(64) march into stored samples (anchor segments in parallel), "
                              "transfer in 32-sample chunks, dual-sweep adjoint, visibility chi2 and seed on the device"},
           "mixed_over_fp64": paths["fp64"]["fwd_adjoint"]["ms_per_step"] / paths["mixed"]["fwd_adjoint"]["ms_per_step"],
           "mixed_error_vs_fp64": {"stokes_over_peak_I": img_err, "gradient_over_max": grad_err},
           "vis_chi2": {p_: runs[p_]["chi2"] for p_ in runs},
           "paths": paths, "gradient_evals_per_s": 1.0 / adj_s,
           "stages": stages,
           "roofline": roofline, "cpu_baseline": cpu, "vis_stage": vis_stage}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
