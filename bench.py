"""Benchmark: ray-samples/s of the fit step (geodesic march + polarized transfer + visibility chi2 + dual-sweep
adjoint) on one MI355X, the mixed FP32 path against the plain FP64 path.

Workload (default: the full fit step the headline metric is quoted on): a = 0.9, theta_o = 163 deg, 512 x 512
screen over +-15 M, 2048 Mino-time samples per ray, 1024 kappa-electron splats (22 rows each) whose centres follow
4 trajectory knots (RK4 in Boyer-Lindquist coordinates, t = -100 .. 200 M), 3 frequencies (86, 230, 345 GHz) x 32
slow-light frames over 100 M, 8192 uv points; seeded synthetic draws (kerrsplat_amd/synthetic.py).

One fit step: the model build (the per-splat constants on the host and their upload), the Recurrence(64) march
of every ray into sample storage, then for each of the 96 (frame, frequency) slices the polarized transfer with
the adjoint records, the slice's visibilities with their chi2 against synthetic data and the adjoint seed of the
weighted residuals (on the device), and the reverse pass that adds the slice's share of the gradient with
respect to all 22 x 1024 splat parameters. A bench step is one slice (the slices cycle through the steps); the
model build and the march happen once per fit step, so each bench step is charged 1/96 of their measured
duration:
    ms_per_step = (slice wall time, averaged over the timed steps) + (model build + march) / 96,
    value       = pixels x samples / ms_per_step  (= the fit step's ray-samples / its duration).
The FP64 path runs the same steps (fewer of them by default: a slice is several times longer); the mixed
path's images and gradient of one slice are compared with the FP64 path's (Stokes error over the peak I,
gradient error over the largest entry), and the mixed path is the headline only within 1e-4 and 1e-3.

Prints one JSON line.
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
    p.add_argument("--steps", type=int, default=5, help="timed slices of the mixed path")
    p.add_argument("--warmup", type=int, default=2, help="untimed slices before them")
    p.add_argument("--fp64-steps", type=int, default=2, help="timed slices of the FP64 path, taken among the mixed path's (at most --steps)")
    p.add_argument("--fp64-warmup", type=int, default=1)
    p.add_argument("--config", type=int, default=4)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    p.add_argument("--paths", default="mixed,fp64",
                   help="the paths to run (a profiling run may take one; the headline needs both)")
    p.add_argument("--no-tune", action="store_true", help="keep the default launch choices (no autotuner)")
    p.add_argument("--launch", default=None,
                   help="fixed launch choices instead of the autotuner (profiling runs): chunk,queue,patch_a,patch_b,"
                        "adjoint_queue_mixed,adjoint_queue_fp64, e.g. 16,1,8,8,0,1")
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


def progress(msg):
    """A line on stderr as each step of the run ends (the JSON result is the only line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def model_parts(syn, cfg_id):
    """The configuration's model as (kind, params[, knots, knot_times]) parts."""
    m = syn.config_model(cfg_id)
    if isinstance(m, list):
        return m
    return [m]


def build_model(K, parts):
    ms = []
    for p in parts:
        kind, P = p[0], p[1]
        if len(p) > 2:
            ms.append(K.Model(P, kind=kind, knots=p[2], knot_times=p[3]))
        else:
            ms.append(K.Model(P, kind=kind))
    return ms[0] if len(ms) == 1 else K.CompositeModel(*ms)


def cpu_baseline(cfg, parts, L, target_s, t_obs, nu, res):
    """The oracle (FP64 C++ restatement of the reference's per-ray CPU path: Recurrence(64) march, transfer and
    dual sweep with every row of every splat, OpenMP over the rays on all host cores) on a bounded sample of the
    same workload at the first slice: one random ray per thread, with as many samples per ray (at most the
    workload's) as fit the target time (the oracle's work per ray-sample does not depend on the sampling: each
    sample visits every splat); fwd+adjoint ray-samples/s."""
    import oracle
    a, th, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["N"]
    from kerrsplat_amd import Camera
    cam = Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    oparts = []
    for p in parts:
        kind, P = p[0], p[1]
        if kind == "kappa":
            hyp = [oracle.kappa_hyp(P[13, i], math.exp(P[14, i])) for i in range(P.shape[1])]
            oparts.append(("kappa", P, hyp, 1e5) + ((p[2], p[3]) if len(p) > 2 else ()))
        elif kind == "powerlaw":
            oparts.append(("powerlaw", P, None, 1e5))
        else:
            oparts.append(tuple(p))
    threads = oracle.num_threads()
    rng = np.random.default_rng(7)
    order = rng.permutation(cam.npix)
    ds = rng.normal(size=(cam.npix, 4))

    def run(n, ns_):
        idx = order[:n]
        t0 = time.perf_counter()
        oracle.model_gradient(a, th, cam.alphas[idx], cam.betas[idx], ns_, oparts, t_obs, nu, L, ds[idx])
        return time.perf_counter() - t0

    n, ns_ = threads, min(N, 32)  # one ray per thread: the oracle runs the rays in parallel
    dt = run(n, ns_)
    progress(f"CPU baseline: {n} rays x {ns_} samples in {dt:.1f} s")
    if dt < 0.5 * target_s and ns_ < N:
        ns_ = int(min(N, max(ns_ + 1, ns_ * target_s / max(dt, 1e-3))))
        dt = run(n, ns_)
        progress(f"CPU baseline: {n} rays x {ns_} samples in {dt:.1f} s")
    return {"value": n * ns_ / dt, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"{n} random rays of the {cam.npix} x {ns_} samples per ray (of {N}) at one (frame, frequency), "
                      f"fwd+adjoint (dual sweep over every splat), FP64 oracle, {dt:.1f} s"}


def visibility_stage(K, syn, cfg, model, L, res=512, nuv=8192, reps=3):
    """The visibility transform's matrix-core product at the largest configuration's size (512 x 512 screen,
    8192 uv points, one frame; the image rendered from the same splats over a fused 64-sample march): HIP-event
    duration of the FP64 GEMM and its rate against the FP64 matrix peak."""
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
        print(json.dumps({"error": "this implementation runs on exactly one GPU; --gpus N > 1 is not supported "
                                   "(the fit step would shard by frame and frequency slices, one GPU per process)"}))
        sys.exit(2)
    if args.config == 2:
        return batched_main(args)
    import kerrsplat_amd as K
    from kerrsplat_amd import api, synthetic as syn

    cfg = syn.CONFIGS[args.config]
    if cfg["kind"] == "thin":
        raise SystemExit("the fit-step bench needs a polarized configuration (1, 2, 3 or 4)")
    a, th, res, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["res"], cfg["N"]
    cam = K.Camera((-cfg["fov"], cfg["fov"]), (-cfg["fov"], cfg["fov"]), res)
    L = syn.length_unit()
    parts = model_parts(syn, args.config)
    times, nus = list(cfg["times"]), list(cfg["nus"])
    slices = [(t, nu) for nu in nus for t in times]            # frame fastest, as the device numbers them
    nsl = len(slices)
    lib, ctx = api._load(), api.context()
    sync = lambda: lib.ks_ctx_sync(ctx)
    print(f"bench: {cfg['name']}: {res}x{res} x {N}, {sum(p[1].shape[1] for p in parts)} splats, {nsl} slices", file=sys.stderr, flush=True)

    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, a, th, marcher=K.Recurrence(64))
    model = build_model(K, parts)
    nuv = cfg["nuv"]
    u, v = syn.uv_disc(args.config, max(nuv, 1))
    psize = syn.pixel_size_rad(-cfg["fov"], cfg["fov"], res)

    # the once-per-fit-step parts, timed on their own: the model build (host constants + upload) and the march
    t0 = time.perf_counter()
    for _ in range(3):
        model = build_model(K, parts)
    sync()
    model_ms = 1e3 * (time.perf_counter() - t0) / 3
    march_ms = []
    for _ in range(3):
        K.march(cache)
        march_ms.append(K.stage_ms("march"))
    sync()
    march_ms = float(np.mean(march_ms))
    once_ms = model_ms + march_ms

    # the transfer's launch choices by the autotuner (on a hashed 1/16 sample of the first slice's tasks, the mixed
    # path with the adjoint records); both paths then run with them
    # the element adjoints' launch form per path (one reverse pass in each form, on the first slice)
    tune = None
    if args.launch:
        ch, qu, pa, pb, aqm, aq64 = (int(x) for x in args.launch.split(","))
        choice = {"chunk": ch, "queue": bool(qu), "patch": (pa, pb), "adjoint_queue": bool(aqm)}
        tune = {"mixed": {"choice": choice, "seconds": 0.0, "table": [], "fixed": True},
                "fp64_adjoint_queue": bool(aq64)}
        progress(f"fixed launch choices: {choice}, FP64 element adjoints queue {bool(aq64)}")
    elif not args.no_tune:
        K.precision(cache, "mixed")
        t0 = time.perf_counter()
        choice, table = K.autotune(cache, model, [slices[0][0]], [slices[0][1]], L, records=True, sample_every=16, adjoint=True)
        tune = {"mixed": {"choice": choice, "seconds": time.perf_counter() - t0,
                          "table": [{"chunk": c_, "patch": list(p_), "queue": q_, "ms": m_} for c_, p_, q_, m_ in table]}}
        progress(f"autotune mixed ({tune['mixed']['seconds']:.1f} s): {choice}")

    # fixed synthetic visibility data: the FP64 model's visibilities of the first slice, perturbed by 5 % and
    # thermal noise at 1 % of the peak, with a different draw for each slice
    K.precision(cache, "fp64")
    rng = np.random.default_rng(20261203)
    vdata, sig = None, None
    if nuv:
        K.gradient_forward(cache, model, [slices[0][0]], [slices[0][1]], L, copy=False)
        V0 = K.visibilities(cache, 1, psize, u, v)[0]
        sig = 0.01 * np.abs(V0[:, 0]).max() * np.ones(V0.shape)
        vdata = [V0 * (1 + 0.05 * rng.normal(size=V0.shape)) + sig * (rng.normal(size=V0.shape) + 1j * rng.normal(size=V0.shape))
                 for _ in range(nsl)]
    ds = np.random.default_rng(20261202).normal(size=(1, 1, cam.npix, 4))     # the image seed when there is no uv set
    dparams = np.zeros(model.grad_shape, order="F")
    STG = ["transfer", "adjoint_steps", "adjoint_elements"] + (["vis"] if nuv else [])

    def slice_step(s, dp, copy=False):
        """One slice: forward pass with the records, the visibility chi2 and seed, the reverse pass."""
        t, nu = slices[s % nsl]
        cube = K.gradient_forward(cache, model, [t], [nu], L, copy=copy)
        chi = None
        if nuv:
            chi = K.vis_chi2_seed(cache, 0, psize, u, v, vdata[s % nsl], sig)
            K.gradient_reverse(cache, model, dp)
        else:
            K.gradient_reverse(cache, model, dp, ds)
        return chi, cube

    def timed(timed_slices, warm_slices):
        """The warmup slices untimed (the first, slice 0, also gives the images and gradient of the error bars), then
        the timed slices, each ended by a stream synchronization; per slice its wall time and each stage's kernel
        time from the HIP events the library records around its launches."""
        g0 = np.zeros(model.grad_shape, order="F")
        tw = time.perf_counter()
        _, cube0 = slice_step(warm_slices[0], g0, copy=True)
        progress(f"warmup slice {warm_slices[0]} done ({time.perf_counter() - tw:.1f} s)")
        for s in warm_slices[1:]:
            slice_step(s, dparams)
            progress(f"warmup slice {s} done ({time.perf_counter() - tw:.1f} s)")
        sync()
        ms = {k: [] for k in STG}
        chis, wall = [], {}
        t0 = time.perf_counter()
        for n, s in enumerate(timed_slices):
            ts = time.perf_counter()
            chis.append(slice_step(s, dparams)[0])
            sync()
            wall[s] = time.perf_counter() - ts
            for k in STG:
                ms[k].append(api.stage_ms(k))
            progress(f"timed slice {n + 1} of {len(timed_slices)} (slice {s}) done ({time.perf_counter() - t0:.1f} s)")
        return (time.perf_counter() - t0) / len(timed_slices), wall, {k: float(np.mean(x)) for k, x in ms.items()}, chis, cube0, g0

    # the mixed path times --steps slices after --warmup; the FP64 path, several times slower, times --fp64-steps of
    # those same slices (spread over them), so that the two are compared on the same work
    n_w = max(args.warmup, 1)
    mixed_timed = list(range(n_w, n_w + args.steps))
    k64 = max(1, min(args.steps, args.fp64_steps))
    fp64_timed = [mixed_timed[int((q + 0.5) * len(mixed_timed) / k64)] for q in range(k64)]
    plan = {"mixed": (mixed_timed, list(range(n_w))), "fp64": (fp64_timed, list(range(max(min(args.warmup, args.fp64_warmup), 1))))}
    runs = {}
    for prec in [p_ for p_ in ("mixed", "fp64") if p_ in args.paths.split(",")]:
        K.precision(cache, prec)
        if tune is not None:
            mixed_choice = tune["mixed"]["choice"]
            K.tuning(cache, chunk=mixed_choice["chunk"], queue=mixed_choice["queue"], patch=mixed_choice["patch"],
                     adjoint_queue=mixed_choice["adjoint_queue"])
            if prec == "fp64" and "fp64_adjoint_queue" in tune:
                K.tuning(cache, adjoint_queue=tune["fp64_adjoint_queue"])
            elif prec == "fp64":
                # the FP64 path keeps the transfer's choices and takes its own element-adjoint form
                t0 = time.perf_counter()
                choice, table = K.autotune(cache, model, [slices[0][0]], [slices[0][1]], L, adjoint=True, transfer=False)
                tune["fp64"] = {"choice": choice, "seconds": time.perf_counter() - t0,
                                "table": [{"chunk": c_, "patch": list(p_), "queue": q_, "ms": m_} for c_, p_, q_, m_ in table]}
                progress(f"autotune fp64 element adjoints ({tune['fp64']['seconds']:.1f} s): {choice}")
        tsl, wsl = plan[prec]
        if prec == "fp64" and "mixed" not in args.paths.split(","):
            tsl = mixed_timed[:k64]
        progress(f"{prec}: {len(wsl)} warmup + {len(tsl)} timed slices")
        wall, per, stage, chis, cube, g = timed(tsl, wsl)
        runs[prec] = {"wall_s": wall, "per_slice_s": per, "stage_ms": stage, "steps": len(tsl), "warmup": len(wsl), "cube": cube,
                      "grad": g, "chi2": chis[-1] if chis else None}
        progress(f"{prec}: {1e3 * wall:.0f} ms per slice {stage}")
    K.precision(cache, "fp64")

    both = "mixed" in runs and "fp64" in runs
    img_err = grad_err = None
    finite = all(bool(np.isfinite(r["grad"]).all() and np.isfinite(r["cube"]).all()) for r in runs.values())
    if both:
        peak = np.abs(runs["fp64"]["cube"][..., 0]).max()
        img_err = float(np.abs(runs["mixed"]["cube"] - runs["fp64"]["cube"]).max() / peak)
        grad_err = float(np.abs(runs["mixed"]["grad"] - runs["fp64"]["grad"]).max() / np.abs(runs["fp64"]["grad"]).max())
        head = "mixed" if (img_err <= 1e-4 and grad_err <= 1e-3 and finite) else "fp64"
    else:
        head = next(iter(runs))         # a profiling run of one path: no error bars, no headline comparison
    per_slice = cam.npix * N
    ms_step = {p_: 1e3 * r["wall_s"] + once_ms / nsl for p_, r in runs.items()}
    # the speed-up on the slices both paths timed (each with its share of the once-per-fit-step work)
    ratio, matched = None, []
    if both:
        matched = sorted(runs["fp64"]["per_slice_s"])
        t64 = sum(runs["fp64"]["per_slice_s"][s] for s in matched) + len(matched) * once_ms / nsl * 1e-3
        t32 = sum(runs["mixed"]["per_slice_s"][s] for s in matched) + len(matched) * once_ms / nsl * 1e-3
        ratio = t64 / t32
    ms_per_step = ms_step[head]
    value = per_slice / (ms_per_step * 1e-3)

    # the dominant kernel's roofline: algorithmic flops per ray-sample from the work model (profiles/work_model.json,
    # the per-ray-sample operation counts of this workload), against the FP32 vector peak on the mixed path and
    # the FP64 vector peak on the FP64 path
    wm_path = os.path.join(ROOT, "profiles", "work_model.json")
    wm = json.load(open(wm_path)).get(cfg["name"]) if os.path.exists(wm_path) else None
    tpl = {"fp64": "double", "mixed": "f32"}
    nsp = sum(p[1].shape[1] for p in parts)
    cull = "true" if 64 <= nsp <= 1024 else "false"          # the transfer's per-tile cull (from 64 to 1024 splats)
    kernel_of = {"transfer": lambda p_: f"ks_chunk_transfer<{model_kind_id(parts)}, {tpl[p_]}, {cull}>",
                 "adjoint_steps": lambda p_: f"ks_adj_segment<{tpl[p_]}>",
                 "adjoint_elements": lambda p_: f"ks_adj_splats<{tpl[p_]}, {adj_pops(parts)}>",
                 "vis": lambda p_: "ks_dgemm<false>"}
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE summary of the workload (rocprofv3 --pmc, scaled by
    # a copy probe); a kernel launched at several grid sizes is keyed name@grid: its largest launch is taken
    tr_file = {1: "traffic_128x128_N1024_32thermal.json", 4: "traffic_512x512_N2048_1024kappa_knots.json"}.get(args.config)
    tr_path = os.path.join(ROOT, "profiles", tr_file) if tr_file else ""
    traffic = {}
    if tr_path and os.path.exists(tr_path):
        for key, v in json.load(open(tr_path))["kernels"].items():
            name = key.split("@")[0]
            if name not in traffic or v["grid"] > traffic[name]["grid"]:
                traffic[name] = v

    def stage_table(prec):
        peak_ = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        out = {}
        for k, t_ms in runs[prec]["stage_ms"].items():
            st = {"kernel": kernel_of[k](prec), "ms": t_ms}
            if st["kernel"] in traffic:
                st["hbm_bytes_per_launch"] = traffic[st["kernel"]]["bytes"]
            ph = (wm or {}).get("phases_fp64_path" if prec == "fp64" and "phases_fp64_path" in (wm or {}) else "phases", {}).get(k)
            if ph and k != "vis":
                achieved = ph["flops_per_ray_sample"] * per_slice / (t_ms * 1e-3) / 1e12
                st.update({"flops_per_ray_sample": ph["flops_per_ray_sample"], "achieved_tflops": achieved, "peak_tflops": peak_,
                           "frac": achieved / peak_})
            out[k] = st
        return out

    stages = {p_: stage_table(p_) for p_ in runs}
    sm = runs[head]["stage_ms"]
    dom = max((k for k in sm if k != "vis"), key=sm.get)
    d = stages[head][dom]
    roofline = None
    if "achieved_tflops" in d:
        roofline = {"bound": "mfma" if dom == "vis" else "valu", "achieved": d["achieved_tflops"], "peak": d["peak_tflops"],
                    "unit": "TFLOP/s", "frac": d["frac"], "traffic": d.get("hbm_bytes_per_launch"), "kernel": d["kernel"],
                    "kernel_ms_per_slice": d["ms"], "flops_per_ray_sample": d["flops_per_ray_sample"],
                    "precision": "fp32" if head == "mixed" else "fp64"}
    paths = {p_: {"value": per_slice / (ms_step[p_] * 1e-3), "unit": "ray-samples/s", "ms_per_step": ms_step[p_],
                  "slice_wall_ms": 1e3 * r["wall_s"], "slice_wall_ms_by_slice": {str(s): 1e3 * t for s, t in r["per_slice_s"].items()},
                  "gradient_evals_per_s": 1.0 / (nsl * ms_step[p_] * 1e-3),
                  "steps": r["steps"], "warmup": r["warmup"], "stage_ms": r["stage_ms"], "vis_chi2_last": r["chi2"]}
             for p_, r in runs.items()}

    vis_stage = visibility_stage(K, syn, cfg, model, L) if nuv else None
    if args.copy_probe:
        copy_probe()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, parts, L, args.cpu_seconds, slices[0][0], slices[0][1], res)
    out = {"metric": METRIC, "value": value, "unit": "ray-samples/s", "n_gpus": 1, "steps": runs[head]["steps"],
           "warmup": runs[head]["warmup"], "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "mixed f32/f64" if head == "mixed" else "f64",
           "data": "synthetic (seeded splat draws, RK4 trajectory knots, uv disc, visibility data from the FP64 model)",
           "config": {"workload": cfg["name"], "a": a, "theta_o_deg": cfg["theta_o_deg"], "screen": [res, res],
                      "samples_per_ray": N, "splats": sum(p[1].shape[1] for p in parts), "kind": cfg["kind"],
                      "knots": len(cfg.get("knot_times", [])), "frames": len(times), "frequencies": len(nus), "uv_points": nuv,
                      "step": "one (frame, frequency) slice of the fit step: transfer with records, visibility chi2 and seed, "
                              "reverse pass; plus 1/%d of the model build and the march" % nsl,
                      "path": ("mixed: FP64 geodesics (Recurrence(64) into stored samples), FP32 transfer and dual-sweep adjoint "
                               "(FP64 step above Faraday depth 10), visibility chi2 and seed on the device (FP64 GEMMs)")
                              if head == "mixed" else
                              "plain FP64: Recurrence(64) march into stored samples, transfer, dual-sweep adjoint, visibility "
                              "chi2 and seed on the device"},
           "mixed_over_fp64": ratio, "mixed_over_fp64_slices": matched,
           "mixed_error_vs_fp64": {"stokes_over_peak_I": img_err, "gradient_over_max": grad_err, "finite": finite},
           "fit_step_s": {p_: nsl * ms_step[p_] * 1e-3 for p_ in runs},
           "once_per_fit_step_ms": {"model_build": model_ms, "march": march_ms},
           "paths": paths, "gradient_evals_per_s": 1.0 / (nsl * ms_per_step * 1e-3),
           "stages": stages, "roofline": roofline, "cpu_baseline": cpu, "vis_stage": vis_stage, "launch_tuning": tune}
    print(json.dumps(out), flush=True)


def batched_main(args):
    """The batched configuration: one forward cube (the polarized transfer of every point of a 256 x 256 screen in
    uneven 32^2-128^2 tiles plus 32 extra screens of U[48, 160]^2, 1024 samples per ray, 128 splats of the three
    populations, 3 frequencies x 16 frames) in one launch of the mixed-precision transfer. The launch is timed as
    one workgroup per task, as the persistent work queue (index order, then the cost order of the previous launch)
    and with the autotuner's choice (timed on a sample of one slice's tasks); a step is one cube with the chosen
    launch. The FP64 path renders the first slice for the error bar of the mixed cube. No adjoint: this
    configuration is a forward render."""
    import kerrsplat_amd as K
    from kerrsplat_amd import api, synthetic as syn
    cfg = syn.CONFIGS[2]
    a, th, N = cfg["a"], math.radians(cfg["theta_o_deg"]), cfg["N"]
    main_cam, splits, extras = syn.config_screens(2)
    tiles, _ = K.screen_tiles(main_cam, splits)
    cam = K.concatenate_screens(*tiles, *extras)
    L = syn.length_unit()
    parts = model_parts(syn, 2)
    model = build_model(K, parts)
    times, nus = list(cfg["times"]), list(cfg["nus"])
    nsl = len(times) * len(nus)
    lib, ctx = api._load(), api.context()
    sync = lambda: lib.ks_ctx_sync(ctx)
    progress(f"{cfg['name']}: {len(tiles)} tiles + {len(extras)} screens = {cam.npix} points x {N}, {model.ns} splats, {nsl} slices")
    cache = K.GeodesicCache(cam, N, store_samples=True)
    K.regenerate(cache, a, th, marcher=K.Recurrence(64))
    sync()
    progress("marched")
    K.precision(cache, "mixed")
    work = cam.npix * N * nsl

    def cube_ms(tag):
        t0 = time.perf_counter()
        K.render_device(cache, model, times, nus, L)
        sync()
        wall = 1e3 * (time.perf_counter() - t0)
        tr = K.stage_ms("transfer")
        progress(f"{tag}: {K.tuning(cache)}: cube {wall:.0f} ms (transfer {tr:.0f} ms)")
        return wall, tr

    runs = {}
    K.tuning(cache, chunk=32, queue=False, patch=(8, 8))
    runs["grid"] = cube_ms("one workgroup per task (warm-up)")
    runs["grid"] = cube_ms("one workgroup per task")
    K.tuning(cache, queue=True)
    runs["queue_index_order"] = cube_ms("work queue, index order")
    runs["queue_cost_order"] = cube_ms("work queue, cost order")
    t0 = time.perf_counter()
    choice, table = K.autotune(cache, model, [times[0]], [nus[1]], L, records=False, sample_every=4)
    tune_s = time.perf_counter() - t0
    progress(f"autotune ({tune_s:.1f} s): {choice}")
    steps = []
    try:
        w, tr = cube_ms("tuned cube (first)")
    except api.KernelError as e:
        # the segment maps of the whole cube at the tuned chunk length did not fit: the next longer chunk
        progress(f"tuned launch failed ({e}); chunk {2 * choice['chunk']}")
        K.tuning(cache, chunk=2 * choice["chunk"])
        choice = K.tuning(cache)
        w, tr = cube_ms("tuned cube (first)")
    for k in range(max(args.warmup, 0) + max(args.steps, 1)):
        w, tr = cube_ms(f"tuned cube {k + 1}")
        if k >= args.warmup:
            steps.append((w, tr))
    if choice["queue"]:
        w, tr = cube_ms("tuned cube (cost order)")
    ms_per_step = float(np.mean([w for w, _ in steps]))
    # the error bar: the first slice on both paths
    img32 = np.array(K.polarized_cube(cache, model, [times[0]], [nus[0]], L))
    K.precision(cache, "fp64")
    img64 = np.array(K.polarized_cube(cache, model, [times[0]], [nus[0]], L))
    sync()
    peak = float(np.abs(img64[..., 0]).max())
    err = float(np.abs(img32 - img64).max() / peak)
    progress(f"mixed vs FP64 on the first slice: {err:.2e} of the peak")
    out = {"metric": "ray-samples/s, forward cube over batched screens (mixed FP32 transfer, FP64 geodesics)", "value": work / (ms_per_step * 1e-3),
           "unit": "ray-samples/s", "n_gpus": 1, "steps": len(steps), "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "mixed f32/f64", "data": "synthetic (seeded draws)",
           "config": {"workload": cfg["name"], "points": cam.npix, "tiles": [[a1 - a0, b1 - b0] for a0, a1, b0, b1 in splits],
                      "extra_screens": [list(c.size) for c in extras], "samples_per_ray": N, "splats": model.ns,
                      "frames": len(times), "frequencies": len(nus), "step": "one forward cube (transfer + composition) of every point, frame and frequency"},
           "launches_ms": {k: {"cube_wall_ms": w, "transfer_ms": tr} for k, (w, tr) in runs.items()},
           "tuned": {"choice": choice, "autotune_s": tune_s, "transfer_ms": float(np.mean([tr for _, tr in steps])),
                     "table": [{"chunk": c_, "patch": list(p_), "queue": q_, "sample_ms": m_} for c_, p_, q_, m_ in table]},
           "mixed_error_vs_fp64_first_slice": err, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def model_kind_id(parts):
    kinds = {"thermal": 1, "powerlaw": 2, "kappa": 3, "knot": 4}
    return 5 if len(parts) > 1 else kinds[parts[0][0]]


def adj_pops(parts):
    """The element-adjoint instantiation of the model's populations (1 thermal, 2 power law, 8 kappa, 4 knots; more
    than one population takes the instantiation of all, 15)."""
    pops, knots = 0, 0
    for p in parts:
        pops |= {"thermal": 1, "knot": 1, "powerlaw": 2, "kappa": 8}[p[0]]
        if p[0] == "knot" or len(p) > 2:
            knots = 4
    return pops | knots if pops in (1, 2, 8) else 15


if __name__ == "__main__":
    main()
