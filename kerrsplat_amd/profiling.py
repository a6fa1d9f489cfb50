"""Summaries of rocprofv3 output for the kernels of this package (text for profiles/, no GPU needed).

    python -m kerrsplat_amd.profiling trace  <kernel_stats.csv>                      > summary.md
    python -m kerrsplat_amd.profiling traffic <fetch counter_collection.csv> <write counter_collection.csv> \
        [--copy-bytes N] [--json out.json]                                           > summary.md

`trace` tabulates the kernel-trace statistics (calls, average and total duration per kernel).
`traffic` turns FETCH_SIZE and WRITE_SIZE (KiB per dispatch) into HBM bytes per launch: each counter is
scaled by what it reads for the runtime's device-to-device copy of a known size (bench.py
--copy-probe copies 256 MiB, so FETCH and WRITE should each read 256 MiB for that dispatch); the
largest copy dispatch is taken as the probe.
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


def counter_means(path, counter):
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            per[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return per


def traffic_summary(fetch_csv, write_csv, copy_bytes=1 << 28):
    fetch, write = counter_means(fetch_csv, "FETCH_SIZE"), counter_means(write_csv, "WRITE_SIZE")
    if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
        raise SystemExit("no copy-probe dispatch in the counter files (run bench.py with --copy-probe)")
    kib = copy_bytes / 1024.0
    f_scale = kib / max(fetch[COPY_KERNEL])
    w_scale = kib / max(write[COPY_KERNEL])
    out = {"copy_probe_bytes": copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        if k == COPY_KERNEL:
            continue
        fk = sum(fetch.get(k, [0.0])) / max(len(fetch.get(k, [])), 1)
        wk = sum(write.get(k, [0.0])) / max(len(write.get(k, [])), 1)
        out["kernels"][k] = {"fetch_bytes": fk * 1024.0 * f_scale, "write_bytes": wk * 1024.0 * w_scale,
                             "bytes": (fk * f_scale + wk * w_scale) * 1024.0, "dispatches": len(fetch.get(k, []))}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats")
    c = sub.add_parser("traffic")
    c.add_argument("fetch")
    c.add_argument("write")
    c.add_argument("--copy-bytes", type=int, default=1 << 28)
    c.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        rows = trace_summary(a.stats)
        print("| kernel | calls | avg (us) | total (ms) | % |\n|---|---|---|---|---|")
        for r in rows:
            print(f"| `{r['kernel']}` | {r['calls']} | {r['avg_us']:.1f} | {r['total_ms']:.2f} | {r['percent']:.2f} |")
    else:
        s = traffic_summary(a.fetch, a.write, a.copy_bytes)
        print(f"copy probe: {s['copy_probe_bytes']} bytes; FETCH_SIZE scale {s['fetch_scale']:.3f}, WRITE_SIZE scale {s['write_scale']:.3f}\n")
        print("| kernel | read (MB) | written (MB) | HBM bytes per launch (MB) |\n|---|---|---|---|")
        for k, v in sorted(s["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
            print(f"| `{k}` | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | {v['bytes'] / 1e6:.1f} |")
        if a.json:
            json.dump(s, open(a.json, "w"), indent=1)


if __name__ == "__main__":
    sys.exit(main())
