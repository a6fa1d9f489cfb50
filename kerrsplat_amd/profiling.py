"""Summaries of rocprofv3 output for the kernels of this package (text for profiles/, no GPU needed).

    python -m kerrsplat_amd.profiling trace  <kernel_stats.csv>                      > summary.md
    python -m kerrsplat_amd.profiling launches <kernel_trace.csv>                     > summary.md
    python -m kerrsplat_amd.profiling valu <counter_collection.csv with SQ_INSTS_VALU, GRBM_GUI_ACTIVE> > summary.md
    python -m kerrsplat_amd.profiling traffic <fetch counter_collection.csv> <write counter_collection.csv> \
        [--copy-bytes N] [--json out.json]                                           > summary.md
    python -m kerrsplat_amd.profiling mix <counter_collection.csv with SQ_INSTS_VALU and the FP32/FP64 counters> > summary.md
    python -m kerrsplat_amd.profiling mfma <counter_collection.csv with SQ_INSTS_VALU_MFMA_MOPS_F64,
        SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE>                                 > summary.md

`trace` tabulates the kernel-trace statistics (calls, average and total duration per kernel).
`launches` does the same from the per-dispatch trace, one row per (kernel, grid size), with the
registers and scratch (spill) bytes per lane the compiler gave the kernel.
`valu` turns SQ_INSTS_VALU and GRBM_GUI_ACTIVE into the VALU issue utilisation of each kernel: a wave64
vector instruction holds a SIMD (16 lanes) for 4 cycles, so an XCD of 32 CUs x 4 SIMDs issues at most 32
wave-instructions per cycle; GRBM_GUI_ACTIVE sums the busy cycles of the 8 XCDs, so the utilisation is
SQ_INSTS_VALU / (32 GRBM_GUI_ACTIVE) (quarter-rate transcendentals make the true issue limit lower still).
`traffic` turns FETCH_SIZE and WRITE_SIZE (KiB per dispatch) into HBM bytes per launch: each counter is
scaled by what it reads for the runtime's device-to-device copy of a known size (bench.py
--copy-probe copies 256 MiB, so FETCH and WRITE should each read 256 MiB for that dispatch); the
largest copy dispatch is taken as the probe.
`mix` splits each kernel's vector instructions by type (FP32 and FP64 FMA, add, multiply and transcendental
counters, as collected). `mfma` turns SQ_INSTS_VALU_MFMA_MOPS_F64 (FP64 matrix operations / 512) into FLOPs per
dispatch and, with the dispatch's duration, a rate against the FP64 matrix peak, and SQ_VALU_MFMA_BUSY_CYCLES into
the matrix-core busy fraction: busy cycles / (GRBM_GUI_ACTIVE of one XCD x 1024 SIMDs), GRBM_GUI_ACTIVE being
summed over the 8 XCDs.
"""
import argparse
import collections
import csv
import json
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    """'void ksd::ks_adj_splats<ksd::f32>(double, ...)' -> 'ks_adj_splats<f32>'."""
    head = name.split("(")[0].replace("void ", "").replace("ksd::", "").strip()
    return head


def trace_summary(stats_csv):
    rows = list(csv.DictReader(open(stats_csv)))
    out = []
    for r in rows:
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
                    "total_ms": float(r["TotalDurationNs"]) / 1e6, "percent": float(r["Percentage"])})
    return out


def launch_summary(trace_csv):
    per = collections.OrderedDict()
    for r in csv.DictReader(open(trace_csv)):
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        key = (short_name(r["Kernel_Name"]), grid)
        d = per.setdefault(key, {"ns": [], "vgpr": int(r["VGPR_Count"]), "agpr": int(r["Accum_VGPR_Count"]),
                                 "scratch": int(r["Scratch_Size"]), "lds": int(r["LDS_Block_Size"])})
        d["ns"].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = []
    for (k, grid), d in per.items():
        out.append({"kernel": k, "grid": grid, "calls": len(d["ns"]), "avg_us": sum(d["ns"]) / len(d["ns"]) / 1e3,
                    "total_ms": sum(d["ns"]) / 1e6, "vgpr": d["vgpr"], "agpr": d["agpr"], "scratch": d["scratch"], "lds": d["lds"]})
    return sorted(out, key=lambda r: -r["total_ms"])


def counter_means(path, counter):
    """Counter values per (kernel, grid size): a kernel launched at several sizes gets one entry per size."""
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            per[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return per


def traffic_summary(fetch_csv, write_csv, copy_bytes=1 << 28):
    """HBM bytes per launch, keyed by kernel name, or 'name@grid' for a kernel launched at several grid sizes."""
    fetch, write = counter_means(fetch_csv, "FETCH_SIZE"), counter_means(write_csv, "WRITE_SIZE")
    fc = [v for (k, _), vs in fetch.items() if k == COPY_KERNEL for v in vs]
    wc = [v for (k, _), vs in write.items() if k == COPY_KERNEL for v in vs]
    if not fc or not wc:
        raise SystemExit("no copy-probe dispatch in the counter files (run bench.py with --copy-probe)")
    kib = copy_bytes / 1024.0
    f_scale, w_scale = kib / max(fc), kib / max(wc)
    out = {"copy_probe_bytes": copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    keys = sorted(set(fetch) | set(write))
    sizes = collections.Counter(k for k, _ in keys)
    for k, grid in keys:
        if k == COPY_KERNEL:
            continue
        fv, wv = fetch.get((k, grid), []), write.get((k, grid), [])
        fk = sum(fv) / max(len(fv), 1)
        wk = sum(wv) / max(len(wv), 1)
        name = k if sizes[k] == 1 else f"{k}@{grid}"
        out["kernels"][name] = {"fetch_bytes": fk * 1024.0 * f_scale, "write_bytes": wk * 1024.0 * w_scale,
                                "bytes": (fk * f_scale + wk * w_scale) * 1024.0, "dispatches": len(fv), "grid": grid}
    return out


def valu_summary(path, top=12):
    """Per (kernel, grid): wave-level VALU and SALU instructions and busy cycles per dispatch, and the VALU issue
    utilisation SQ_INSTS_VALU / (32 GRBM_GUI_ACTIVE), heaviest first."""
    valu, salu, busy = (counter_means(path, c) for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "GRBM_GUI_ACTIVE"))
    rows = []
    for key, v in valu.items():
        b = busy.get(key)
        if not b:
            continue
        iv, ib = sum(v) / len(v), sum(b) / len(b)
        isl = salu.get(key, [0.0])
        rows.append({"kernel": key[0], "grid": key[1], "dispatches": len(v), "valu": iv, "salu": sum(isl) / len(isl), "cycles": ib,
                     "util": iv / (32.0 * ib) if ib > 0 else 0.0})
    rows.sort(key=lambda r: -r["valu"])
    return rows[:top]


def mix_summary(path, top=8):
    """Per (kernel, grid), heaviest first: SQ_INSTS_VALU and every FP32 / FP64 instruction counter present, per
    dispatch (wave instructions)."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        per[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for d in per.values() for c in d if c.startswith("SQ_INSTS_VALU_")})
    rows = []
    for key, d in per.items():
        if "SQ_INSTS_VALU" not in d:
            continue
        mean = {c: sum(v) / len(v) for c, v in d.items()}
        rows.append((key, len(d["SQ_INSTS_VALU"]), mean))
    rows.sort(key=lambda r: -r[2]["SQ_INSTS_VALU"])
    return names, rows[:top]


MFMA_F64_PEAK = 78.6e12


def mfma_summary(path):
    """Per (kernel, grid) with FP64 matrix operations: FLOPs per dispatch (MOPS x 512), duration, rate, fraction of
    the FP64 matrix peak, and the matrix-core busy fraction."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        key = (short_name(r["Kernel_Name"]), int(r["Grid_Size"]))
        per[key][r["Counter_Name"]].append(float(r["Counter_Value"]))
        if r["Counter_Name"] == "GRBM_GUI_ACTIVE":
            dur[key].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = []
    for key, d in per.items():
        mops = d.get("SQ_INSTS_VALU_MFMA_MOPS_F64")
        if not mops or sum(mops) == 0 or not dur[key]:
            continue
        flops = 512.0 * sum(mops) / len(mops)
        ns = sum(dur[key]) / len(dur[key])
        busy = sum(d["SQ_VALU_MFMA_BUSY_CYCLES"]) / len(d["SQ_VALU_MFMA_BUSY_CYCLES"]) if "SQ_VALU_MFMA_BUSY_CYCLES" in d else None
        gui = sum(d["GRBM_GUI_ACTIVE"]) / len(d["GRBM_GUI_ACTIVE"]) / 8.0 if "GRBM_GUI_ACTIVE" in d else None
        rows.append({"kernel": key[0], "grid": key[1], "dispatches": len(mops), "flops": flops, "us": ns / 1e3,
                     "tflops": flops / ns / 1e3, "frac": flops / (ns * 1e-9) / MFMA_F64_PEAK,
                     "busy": busy / (gui * 1024.0) if busy is not None and gui else None})
    rows.sort(key=lambda r: -r["flops"])
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats")
    l_ = sub.add_parser("launches")
    l_.add_argument("trace")
    v = sub.add_parser("valu")
    v.add_argument("counters")
    c = sub.add_parser("traffic")
    c.add_argument("fetch")
    c.add_argument("write")
    c.add_argument("--copy-bytes", type=int, default=1 << 28)
    c.add_argument("--json")
    x = sub.add_parser("mix")
    x.add_argument("counters")
    m = sub.add_parser("mfma")
    m.add_argument("counters")
    a = ap.parse_args(argv)
    if a.cmd == "mix":
        names, rows = mix_summary(a.counters)
        short = [n.replace("SQ_INSTS_VALU_", "") for n in names]
        print("| kernel | grid (threads) | dispatches | VALU (wave instr.) | " + " | ".join(short) + " |\n|" + "---|" * (4 + len(names)))
        for key, nd, mean in rows:
            print(f"| `{key[0]}` | {key[1]} | {nd} | {mean['SQ_INSTS_VALU']:.3g} | " + " | ".join(f"{mean.get(n, 0.0):.3g}" for n in names) + " |")
        return
    if a.cmd == "mfma":
        print("| kernel | grid (threads) | dispatches | FP64 matrix FLOPs per dispatch | avg (us) | TFLOP/s | of the FP64 matrix peak | MFMA busy |\n"
              "|---|---|---|---|---|---|---|---|")
        for r in mfma_summary(a.counters):
            busy = f"{100 * r['busy']:.0f}%" if r["busy"] is not None else "-"
            print(f"| `{r['kernel']}` | {r['grid']} | {r['dispatches']} | {r['flops']:.4g} | {r['us']:.1f} | {r['tflops']:.1f} | "
                  f"{100 * r['frac']:.0f}% | {busy} |")
        return
    if a.cmd == "valu":
        print("| kernel | grid (threads) | VALU instr. (wave) | SALU instr. (wave) | busy cycles (8 XCDs) | VALU issue utilisation |\n"
              "|---|---|---|---|---|---|")
        for r in valu_summary(a.counters):
            print(f"| `{r['kernel']}` | {r['grid']} | {r['valu']:.3g} | {r['salu']:.3g} | {r['cycles']:.3g} | {100 * r['util']:.0f}% |")
        return
    if a.cmd == "trace":
        rows = trace_summary(a.stats)
        print("| kernel | calls | avg (us) | total (ms) | % |\n|---|---|---|---|---|")
        for r in rows:
            print(f"| `{r['kernel']}` | {r['calls']} | {r['avg_us']:.1f} | {r['total_ms']:.2f} | {r['percent']:.2f} |")
    elif a.cmd == "launches":
        print("| kernel | grid (threads) | calls | avg (us) | total (ms) | VGPR | AGPR | scratch (B/lane) | LDS (B) |\n"
              "|---|---|---|---|---|---|---|---|---|")
        for r in launch_summary(a.trace):
            print(f"| `{r['kernel']}` | {r['grid']} | {r['calls']} | {r['avg_us']:.1f} | {r['total_ms']:.2f} | {r['vgpr']} | "
                  f"{r['agpr']} | {r['scratch']} | {r['lds']} |")
    else:
        s = traffic_summary(a.fetch, a.write, a.copy_bytes)
        print(f"copy probe: {s['copy_probe_bytes']} bytes; FETCH_SIZE scale {s['fetch_scale']:.3f}, WRITE_SIZE scale {s['write_scale']:.3f}\n")
        print("| kernel | grid (threads) | read (MB) | written (MB) | HBM bytes per launch (MB) |\n|---|---|---|---|---|")
        for k, v in sorted(s["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
            print(f"| `{k.split('@')[0]}` | {v['grid']} | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | {v['bytes'] / 1e6:.1f} |")
        if a.json:
            json.dump(s, open(a.json, "w"), indent=1)


if __name__ == "__main__":
    sys.exit(main())
