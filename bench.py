"""Benchmark: ray-samples/s of the fused geodesic march + polarized transfer on one MI355X.

Workload (the configuration the headline metric is quoted on that fits one GPU): a = 0.5,
theta_o = 60 deg, 128 x 128 screen over +-15 M, 1024 Mino-time samples per ray, 32 thermal splats,
full IQUV at 230 GHz, visibilities on 512 uv points (seeded synthetic draws). One step = one
polarized render of the screen through the fused Recurrence(64) march (K2 + transfer) plus the
visibilities of the image, with the per-pixel constants (K0/K1) regenerated once beforehand and
the splat parameters resident on the device.

Prints one JSON line; value = ray-samples per second (pixels x samples x frames x frequencies / s).
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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    return p.parse_args()


def cpu_baseline(cfg, P, L, target_s):
    """The oracle (FP64 C++ restatement of the reference's per-ray CPU path, OpenMP over all host
    cores) on a bounded, strided sample of the same rays; ray-samples/s."""
    import oracle
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    from kerrsplat_amd import Camera
    cam = Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    threads = oracle.num_threads()
    rng = np.random.default_rng(7)
    order = rng.permutation(cam.npix)
    n = 64
    t0 = time.perf_counter()
    oracle.polarized_image(a, th, cam.alphas[order[:n]], cam.betas[order[:n]], N, P, 0.0, cfg["nus"][0], L)
    dt = time.perf_counter() - t0
    n2 = int(min(cam.npix, max(n, n * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    oracle.polarized_image(a, th, cam.alphas[order[:n2]], cam.betas[order[:n2]], N, P, 0.0, cfg["nus"][0], L)
    dt2 = time.perf_counter() - t0
    return {"value": n2 * N / dt2, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"{n2} of {cam.npix} rays (random subset) x {N} samples, FP64 oracle, {dt2:.1f} s"}


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
    cache = K.GeodesicCache(cam, N, store_samples=False)
    K.regenerate(cache, a, th, marcher=K.Fused(64))
    model = K.Model(P, kind="thermal")
    u, v = syn.uv_disc(args.config, cfg["nuv"])
    psize = syn.pixel_size_rad(-cfg["fov"], cfg["fov"], res)
    lib, ctx = api._load(), api.context()
    times, nus = cfg["times"], cfg["nus"]

    def step():
        K.render_device(cache, model, times, nus, L)
        if cfg["nuv"]:
            K.visibilities(cache, 1, psize, u, v)

    for _ in range(args.warmup):
        step()
    lib.ks_ctx_sync(ctx)
    # the timed region: K steps, synchronized on both sides; the dominant kernel's duration per step
    # (the segment march + transfer, ks_seg_polarized) from the HIP events the library records around
    # its launch on the context's stream, read after the timed region
    wall0 = time.perf_counter()
    kms = []
    for _ in range(args.steps):
        step()
        ms = api.ctypes.c_float()
        if lib.ks_last_kernel_ms(ctx, api.ctypes.byref(ms)) == 0:
            kms.append(ms.value)
    lib.ks_ctx_sync(ctx)
    wall = time.perf_counter() - wall0
    ms_per_step = 1e3 * wall / args.steps
    work = cam.npix * N * len(times) * len(nus)
    value = work / (wall / args.steps)
    kernel_ms = float(np.mean(kms))

    roofline = None
    wm_path = os.path.join(ROOT, "profiles", "work_model.json")
    if os.path.exists(wm_path):
        wm = json.load(open(wm_path)).get(cfg["name"])
        if wm:
            flops = wm["flops_per_ray_sample"] * work
            achieved = flops / (kernel_ms * 1e-3) / 1e12
            roofline = {"bound": "valu-fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": wm.get("hbm_bytes_per_launch"),
                        "kernel": "ks_seg_polarized", "kernel_ms": kernel_ms,
                        "flops_per_ray_sample": wm["flops_per_ray_sample"]}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, P, L, args.cpu_seconds)
    out = {"metric": METRIC, "value": value, "unit": "ray-samples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded splat draws and uv disc)",
           "config": {"workload": cfg["name"], "a": a, "theta_o_deg": cfg["theta_o_deg"], "screen": [res, res],
                      "samples_per_ray": N, "splats": cfg["ns"], "kind": cfg["kind"], "frames": len(times),
                      "frequencies": len(nus), "uv_points": cfg["nuv"], "path": "plain FP64 forward (Recurrence(64) fused, anchor-segment parallel)"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
