"""Profiling helpers for the BP path on MI355X (rocprofv3 output -> summaries).

    python -m mrfmap_amd.profiling copy
        One plain device-to-device copy of 1 GiB (torch), the reference
        workload for calibrating the FETCH_SIZE / WRITE_SIZE counters: an
        ideal copy fetches and writes exactly 1 GiB.
    python -m mrfmap_amd.profiling stream [--gib N] [--reps R]
        Achievable HBM bandwidth of plain torch kernels on N GiB buffers:
        copy and add (read and write N GiB), fill (writes N GiB) and sum
        (reads N GiB), best of R timed repetitions (HIP events), in GB/s.
    python -m mrfmap_amd.profiling summarize DIR [--calib CALDIR] [--kernels A,B] [--first N]
        Markdown summary of rocprofv3 runs kept under DIR: trace/ (from
        --kernel-trace --stats), fetch/ and write/ (from --pmc FETCH_SIZE and
        --pmc WRITE_SIZE, csv output).  With --calib, CALDIR/fetch and
        CALDIR/write hold the same counters for the copy above; the counters
        of every kernel are then scaled by 1 GiB / (counter of the copy).
        Also prints one JSON line with the calibrated HBM bytes per launch of
        the kernel families in --kernels (default: the front-sum and window
        kernels of one camera pass), summed.  With --first N, each kernel's
        figures use its first N dispatches only (from the per-dispatch trace),
        e.g. the 20 x 13 camera passes of the bench's first timed call.
    python -m mrfmap_amd.profiling geometry DIR
        Markdown table of the geometry build in the kernel trace under DIR:
        every dispatch before the first front-sum kernel (activation, lists,
        pairs, window segments), summed per kernel.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

GIB = 1 << 30


def short_name(name):
    """Kernel name without namespaces and argument lists; template arguments kept."""
    n = name.replace("mrf::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    n = n.replace("void ", "")
    n = re.sub(r"\(.*$", "", n)
    n = n.replace("rocprim::ROCPRIM_400200_NS::detail::", "rocprim::")
    return n[:70]


def _read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d, suffix):
    for root, _, files in os.walk(d):
        for fn in files:
            if fn.endswith(suffix):
                return os.path.join(root, fn)
    return None


def counters(d, first=None):
    """{kernel short name: [values]} of the single counter collected under d
    (with first: each kernel's first `first` dispatches)."""
    path = _find(d, "counter_collection.csv")
    if path is None:
        return {}, None
    out = collections.defaultdict(list)
    cname = None
    for r in sorted(_read_csv(path), key=lambda r: int(r["Dispatch_Id"])):
        v = out[short_name(r["Kernel_Name"])]
        if not first or len(v) < first:
            v.append(float(r["Counter_Value"]))
        cname = r["Counter_Name"]
    return out, cname


def kernel_stats(d, first=None):
    """rocprofv3's per-kernel statistics; with first, recomputed from the
    per-dispatch trace over each kernel's first `first` dispatches."""
    if not first:
        path = _find(d, "kernel_stats.csv")
        return _read_csv(path) if path else []
    path = _find(d, "kernel_trace.csv")
    if path is None:
        return []
    durs = collections.OrderedDict()
    for r in sorted(_read_csv(path), key=lambda r: int(r["Dispatch_Id"])):
        v = durs.setdefault(r["Kernel_Name"], [])
        if len(v) < first:
            v.append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    total = sum(sum(v) for v in durs.values()) or 1
    rows = [{"Name": k, "Calls": str(len(v)), "TotalDurationNs": str(sum(v)), "AverageNs": str(sum(v) / len(v)),
             "Percentage": str(100.0 * sum(v) / total)} for k, v in durs.items()]
    return sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))


def geometry(d, top=24):
    """Per-kernel totals of the dispatches before the first k_front_sums (the
    geometry build of the traced run), as markdown."""
    path = _find(d, "kernel_trace.csv")
    if path is None:
        return "no kernel trace under " + d
    rows = sorted(_read_csv(path), key=lambda r: int(r["Start_Timestamp"]))
    first = next((i for i, r in enumerate(rows) if "k_front_sums" in r["Kernel_Name"]), len(rows))
    rows = rows[:first]
    if not rows:
        return "no geometry dispatches"
    agg = collections.OrderedDict()
    for r in rows:
        n = short_name(r["Kernel_Name"])
        n = re.sub(r"rocprim::trampoline_kernel<.*", "rocprim (radix sort and scans)", n)
        n = re.sub(r"rocprim::init_lookback.*", "rocprim (radix sort and scans)", n)
        n = re.sub(r"rocprim::.*", "rocprim (radix sort and scans)", n)
        a = agg.setdefault(n, [0, 0])
        a[0] += 1
        a[1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    total = sum(v[1] for v in agg.values())
    span = int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])
    lines = [f"kernel time {total / 1e6:.2f} ms over {len(rows)} dispatches "
             f"(first to last dispatch {span / 1e6:.2f} ms under the tracer)", "",
             "| kernel | calls | total ms | avg ms |", "|---|---:|---:|---:|"]
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        lines.append(f"| {n} | {c} | {t / 1e6:.2f} | {t / c / 1e6:.3f} |")
    return "\n".join(lines)


def calibration(caldir):
    """Scale factors (bytes per counted KiB) for FETCH_SIZE and WRITE_SIZE from the 1 GiB copy."""
    fac = {}
    for kind in ("fetch", "write"):
        c, _ = counters(os.path.join(caldir, kind))
        vals = [v for k, vs in c.items() if "copy" in k.lower() or "elementwise" in k.lower() for v in vs]
        if not vals:
            vals = [max(v for vs in c.values() for v in vs)] if c else []
        if vals:
            fac[kind] = GIB / max(vals)
    return fac


def summarize(d, caldir=None, kernels=("k_front_sums", "k_ray_window"), top=16, first=None):
    lines = []
    parts = []
    stats = kernel_stats(os.path.join(d, "trace"), first)
    if stats:
        lines.append("| kernel | calls | avg ms | total ms | % |")
        lines.append("|---|---:|---:|---:|---:|")
        for r in stats[:top]:
            lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                         f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f} |")
    fac = calibration(caldir) if caldir else {}
    fetch, _ = counters(os.path.join(d, "fetch"), first)
    write, _ = counters(os.path.join(d, "write"), first)
    traffic = None
    if fetch or write:
        lines.append("")
        if fac:
            lines.append(f"Calibration (1 GiB copy): FETCH_SIZE x {fac.get('fetch', 1) / 1024:.3f}, "
                         f"WRITE_SIZE x {fac.get('write', 1) / 1024:.3f} (counted KiB -> KiB)")
            lines.append("")
        lines.append("| kernel | launches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | calibrated GB/launch |")
        lines.append("|---|---:|---:|---:|---:|")
        names = sorted(set(fetch) | set(write), key=lambda k: -(sum(fetch.get(k, [0])) / max(1, len(fetch.get(k, [1])))))
        for k in names[:top]:
            f = fetch.get(k, [])
            w = write.get(k, [])
            fa = sum(f) / len(f) if f else 0.0
            wa = sum(w) / len(w) if w else 0.0
            gb = (fa * fac.get("fetch", 1024) + wa * fac.get("write", 1024)) / 1e9
            lines.append(f"| {k} | {max(len(f), len(w))} | {fa:.4g} | {wa:.4g} | {gb:.3f} |")
        # per kernel family (the variant launched most often: the schedule in use)
        for pat in kernels:
            best = None
            for k in names:
                f = fetch.get(k, [])
                w = write.get(k, [])
                if pat in k and (best is None or len(f) > best["launches"]):
                    fa = sum(f) / len(f) if f else 0.0
                    wa = sum(w) / len(w) if w else 0.0
                    best = {"kernel": k, "launches": len(f), "fetch_kib": fa, "write_kib": wa,
                            "bytes": fa * fac.get("fetch", 1024) + wa * fac.get("write", 1024)}
            if best:
                parts.append(best)
    if parts:
        traffic = {"kernels": parts, "fetch_factor": fac.get("fetch", 1024) / 1024,
                   "write_factor": fac.get("write", 1024) / 1024, "bytes": sum(p["bytes"] for p in parts)}
    return "\n".join(lines), traffic


def copy_once():
    import torch

    n = GIB // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {n * 4} bytes")


def stream(gib=4, reps=10):
    import torch

    n = gib * GIB // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    out = {}
    cases = (("copy", lambda: dst.copy_(src), 2 * n * 4), ("add", lambda: torch.add(src, 1.0, out=dst), 2 * n * 4),
             ("fill", lambda: dst.fill_(2.0), n * 4), ("sum", lambda: src.sum(), n * 4))
    for name, fn, nbytes in cases:
        fn()
        best = None
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        out[name] = {"ms": best, "GB/s": nbytes / (best * 1e-3) / 1e9}
    print(json.dumps({"buffer_gib": gib, "reps": reps, **out}))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("copy")
    st = sub.add_parser("stream")
    st.add_argument("--gib", type=int, default=4)
    st.add_argument("--reps", type=int, default=10)
    s = sub.add_parser("summarize")
    s.add_argument("dir")
    s.add_argument("--calib")
    s.add_argument("--kernels", default="k_front_sums,k_ray_window",
                   help="comma-separated kernel families whose per-launch bytes are summed")
    s.add_argument("--first", type=int, default=None, help="each kernel's first N dispatches only")
    g = sub.add_parser("geometry")
    g.add_argument("dir")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_once()
        return 0
    if a.cmd == "stream":
        stream(a.gib, a.reps)
        return 0
    if a.cmd == "geometry":
        print(geometry(a.dir))
        return 0
    text, traffic = summarize(a.dir, a.calib, tuple(a.kernels.split(",")), first=a.first)
    print(text)
    if traffic:
        print()
        print(json.dumps(traffic))
    return 0


if __name__ == "__main__":
    sys.exit(main())
