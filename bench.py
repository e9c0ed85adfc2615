#!/usr/bin/env python3
"""Benchmark of the MRFMap BP inference path on one MI355X.

Workload "real": 13 synthetic keyframes of 640x480 (f=525, c=(319.5,239.5)),
1 cm voxels in a 4x4x3 m room with 12 boxes and 4 spheres (seed 2), full-ray
brick activation, FP32 messages.  A step is one BP iteration: every camera's
ray-message pass + epilogue + the FP64 residual, with all keyframes, lists
and map state resident in HBM.

Prints one JSON line: updates/s over the timed steps; the time-to-converged
map -- from "keyframes uploaded" (traversal included) to a map within 1e-3
max-abs marginal error of plain FP32 BP iterated to a residual below 3e-5
(the FP64 map is out of the CPU oracle's reach at this size; FP32 and FP64
fixed points agree to ~1e-5 where the oracle can run) -- for plain FP32 BP,
for the learned warm start (mrfmap_amd/data/warm_mlp.npz) with FP32
messages and with FP16 ray->voxel messages; the roofline of the ray-message
kernels and the CPU baseline (the FP64 oracle on a bounded sample of the
same workload, one host core).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="real")
    ap.add_argument("--msg", default="f32", choices=["f32", "f16", "bf16", "mixed-f16", "mixed-bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tol", type=float, default=3e-5, help="residual of the reference run")
    ap.add_argument("--no-converge", action="store_true", help="skip the time-to-converged run (profiling)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the steps in the other launch mode (profiling: with --warmup 0 and --no-converge "
                         "the ray kernels of the trace are exactly those of the timed steps)")
    ap.add_argument("--legs", default="plain,warm,warm-f16",
                    help="time-to-converged runs: plain (FP32 BP), warm (learned warm start, FP32), "
                         "warm-f16 (learned warm start, FP16 ray->voxel messages)")
    ap.add_argument("--launch", default="graph", choices=["timed", "graph"],
                    help="launch mode of the timed steps that give value: graph replays (the library's "
                         "default) or kernels one by one with per-kernel HIP events; the other mode is timed "
                         "afterwards, and the roofline's kernel durations come from the one-by-one steps")
    return ap.parse_args()


def cpu_baseline(wl):
    """FP64 scalar oracle on keyframes 0 and 6 of the same workload, every 4th
    pixel row; one BP iteration, timed on one host core."""
    import oracle

    p, sub = _sample(wl)
    o = oracle.Oracle(p.K, p.rows, p.cols, p.res, p.prior, p.dims, p.const_sigmasq_poly, p.const_bias_poly,
                      lookup_n=p.lookup_n, activation=p.activation)
    for pose, K, d in sub:
        o.add_camera(pose, d, K)
    o.prepare()
    t0 = time.perf_counter()
    st = o.perform_inference(1)
    dt = time.perf_counter() - t0
    return {"value": st["updates_per_iter"] / dt, "unit": "ray-msg updates/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle BP, 1 iteration over keyframes 0 and 6 of '{wl}' with every 4th pixel row "
                      f"({st['updates_per_iter']} updates, {dt:.2f} s)"}


def _sample(wl, keyframes=(0, 6), row_step=4):
    """Some keyframes of the workload with every row_step-th pixel row kept."""
    import numpy as np

    from mrfmap_amd import synth

    p, frames, _ = synth.make_workload(wl, frames=13)
    sub = []
    for k in keyframes:
        pose, K, d = frames[k]
        dd = np.zeros_like(d)
        dd[::row_step] = d[::row_step]
        sub.append((pose, K, dd))
    return p, sub


def cpu_comparator(wl):
    """The reference's CPU comparator path, GVDBOctomapWrapper insertion
    (GVDBOctomapWrapper.cpp:31-43): an OctoMap-equivalent restatement (OcTree
    at the map resolution, hit 0.7 / miss 0.4, clamping 0.0001 / 0.9999,
    bounding box = the map) inserting each keyframe's point cloud from its
    camera centre, one host core, on keyframe 0 with every 8th pixel row."""
    import numpy as np

    import oracle

    p, sub = _sample(wl, keyframes=(0,), row_step=8)
    tree = oracle.OctomapEq(p.res, p.dims, p.occ_thresh)
    clouds = []
    for pose, K, d in sub:
        v, u = np.nonzero(d > 0)
        z = d[v, u].astype(np.float64)
        cam = np.stack([(u - K[0, 2]) / K[0, 0] * z, (v - K[1, 2]) / K[1, 1] * z, z], axis=1)
        world = cam @ pose[:3, :3].astype(np.float64).T + pose[:3, 3].astype(np.float64)
        clouds.append((world.astype(np.float32), pose[:3, 3].astype(np.float64)))
    t0 = time.perf_counter()
    for pts, origin in clouds:
        tree.insert_point_cloud(pts, origin)
    dt = time.perf_counter() - t0
    n = sum(len(c[0]) for c in clouds)
    return {"value": n / dt, "unit": "rays inserted/s", "cores": 1, "kind": "port",
            "sample": f"OctoMap-equivalent insertion of keyframe 0 of '{wl}' with every 8th pixel row "
                      f"({n} rays, {dt:.2f} s)"}


def time_to_converged(p, frames, ref, which):
    """Plain vs warm-started BP to a map within 1e-3 of ``ref`` (fresh
    contexts, one at a time); ``which`` lists the legs to run."""
    from mrfmap_amd import _lib
    from mrfmap_amd import warmstart as W

    params, meta = W.load_weights()
    keep = ("ms", "geometry_ms", "iters", "error", "reached", "warm", "warm_ms")
    legs = {}
    if "plain" in which:
        legs["plain_f32"] = W.time_to_error(p, frames, ref, 1e-3, max_iters=300)
    if "warm" in which:
        legs["warm_f32"] = W.time_to_error(p, frames, ref, 1e-3, params=params, warm_after=meta["warm_after"],
                                           max_iters=300)
    if "warm-f16" in which:
        msg0 = p.msg_dtype
        try:
            p.msg_dtype = _lib.MSG_MIXED_F16
            legs["warm_mixed_f16"] = W.time_to_error(p, frames, ref, 1e-3, params=params,
                                                     warm_after=meta["warm_after"], max_iters=300)
        finally:
            p.msg_dtype = msg0
    out = {k: {f: v[f] for f in keep} for k, v in legs.items()}
    for k, v in out.items():
        print(f"[bench] time-to-1e-3 {k}: {v['iters']} iters {v['ms']:.1f} ms reached={v['reached']} "
              f"error {v['error']:.1e}", file=sys.stderr, flush=True)
    best = min((v["ms"] for k, v in out.items() if k != "plain_f32" and v["reached"]), default=None)
    out["speedup_warm_vs_plain"] = (out["plain_f32"]["ms"] / best) if (best and "plain_f32" in out) else None
    out["warm_network"] = {"file": "mrfmap_amd/data/warm_mlp.npz", "warm_after": meta["warm_after"],
                           "training_seeds": meta["seeds"], "samples": meta["samples"]}
    return out


def measured_traffic(a):
    """HBM bytes per ray-kernel launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated on a
    1 GiB copy), or None when it does not cover this configuration."""
    if a.workload != "real" or a.msg != "f32":
        return None, None
    path = os.path.join(ROOT, "profiles", "ray_kernels_real_13kf_f32_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["bytes"], os.path.relpath(path, ROOT)
    except (OSError, ValueError, KeyError):
        return None, None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): this build measures one GPU's share; "
              "keyframes would shard across GPUs with one belief all-reduce per iteration, which is not built",
              file=sys.stderr)
        sys.exit(2)
    from mrfmap_amd import GVDBInference, _lib, synth

    msg = {"f32": _lib.MSG_F32, "f16": _lib.MSG_F16, "bf16": _lib.MSG_BF16, "mixed-f16": _lib.MSG_MIXED_F16,
           "mixed-bf16": _lib.MSG_MIXED_BF16}[a.msg]
    t0 = time.perf_counter()
    p, frames, _ = synth.make_workload(a.workload, msg_dtype=msg)
    t_gen = time.perf_counter() - t0
    print(f"[bench] generated {len(frames)} keyframes in {t_gen:.1f}s", file=sys.stderr, flush=True)

    def build():
        g = GVDBInference(params=p)
        for pose, K, depth in frames:
            g.add_camera_with_depth(pose, depth, K)
        return g

    # reference map: plain BP from a fresh map to a residual below tol; its
    # context then serves the timed steps and is closed before the
    # time-to-converged runs (a 64-keyframe context fills half the HBM)
    g = build()
    ttc = None
    if a.no_converge:
        t0 = time.perf_counter()
        g.prepare()
        conv = {"iters": 0, "residual": None, "t_total_ms": None,
                "t_geometry_ms": (time.perf_counter() - t0) * 1e3}
    else:
        conv = g.perform_inference(max_iters=400, tol=a.tol)
        print(f"[bench] reference: {conv['iters']} iters residual {conv['residual']:.2e} "
              f"total {conv['t_total_ms']:.1f} ms (geometry {conv['t_geometry_ms']:.1f} ms)", file=sys.stderr,
              flush=True)
    ref = g.get_alphas() if (not a.no_converge and msg == _lib.MSG_F32) else None

    # throughput: W warmup steps then exactly K timed steps on resident data
    modes = {"timed": _lib.LAUNCH_TIMED, "graph": _lib.LAUNCH_GRAPH}
    g.set_launch_mode(modes[a.launch])
    if a.warmup > 0:
        g.perform_inference(max_iters=a.warmup)
    t1 = time.perf_counter()
    st = g.perform_inference(max_iters=a.steps)
    dt = time.perf_counter() - t1
    assert st["iters"] == a.steps
    # the same steps in the other launch mode, for comparison
    other = "graph" if a.launch == "timed" else "timed"
    dt_other = None
    if not a.no_compare:
        g.set_launch_mode(modes[other])
        g.perform_inference(max_iters=max(1, a.warmup))
        t2 = time.perf_counter()
        st_other = g.perform_inference(max_iters=a.steps)
        dt_other = time.perf_counter() - t2
    elif a.launch != "timed":
        sys.exit("--no-compare needs --launch timed: the kernel durations come from the timed mode")
    timed = st if a.launch == "timed" else st_other
    g.close()  # one context at a time: the convergence runs below build their own
    if ref is not None:
        ttc = time_to_converged(p, frames, ref, a.legs.split(","))
    n_upd = st["updates_per_iter"]
    ncam = len(frames)
    counts = {}
    for _, K, depth in frames:
        key = f"{depth.shape[1]}x{depth.shape[0]} f={float(K[0][0])}"
        counts[key] = counts.get(key, 0) + 1
    shapes = ", ".join(f"{n} x {k}" for k, n in sorted(counts.items()))
    value = n_upd * a.steps / dt
    launches = a.steps * ncam
    avg_launch_ms = timed["t_ray_kernel_ms"] / launches
    # ray kernel, algorithmic bytes: every incidence reads its 8 B record and
    # its pair's voxel->ray message; a window incidence also reads its 4 B
    # segment position and writes its ray->voxel message.  Front records of
    # chunks whose sums stop changing are read in the first two iterations of
    # a call only; the average over the K timed launches of a camera counts them so.
    in_b = 4 if msg in (_lib.MSG_F32, _lib.MSG_MIXED_F16, _lib.MSG_MIXED_BF16) else 2
    out_b = 4 if msg == _lib.MSG_F32 else 2
    n_win = st["window_updates_per_iter"]
    n_static = st["static_front_updates"]
    front_b, window_b = 8 + in_b, 8 + in_b + 4 + out_b
    front_reads = (n_upd - n_win - n_static) + n_static * min(a.steps, 2) / a.steps
    bytes_per_launch = (front_reads * front_b + n_win * window_b) / ncam
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(a)
    other_s = f"{dt_other * 1e3:.1f} ms" if dt_other is not None else "not run"
    print(f"[bench] {a.steps} steps ({a.launch}) in {dt * 1e3:.1f} ms, ({other}) {other_s}; "
          f"ray kernels {timed['t_ray_kernel_ms']:.1f} ms, epilogue {timed['t_epilogue_ms']:.1f} ms",
          file=sys.stderr, flush=True)

    cpu = comp = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a.workload)
        print(f"[bench] cpu baseline {cpu['value']:.3e} updates/s", file=sys.stderr, flush=True)
        comp = cpu_comparator(a.workload)
        print(f"[bench] cpu comparator {comp['value']:.3e} rays/s", file=sys.stderr, flush=True)

    out = {
        "metric": "ray-msg updates/s; time-to-converged map (\u22641e-3 vs FP64), warm vs plain BP",
        "value": value,
        "unit": "ray-msg updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.msg,
        "data": "synthetic",
        "config": {
            "workload": f"{a.workload}: {ncam} keyframes ({shapes}), "
                        f"{p.res * 100:.0f} cm voxels, {list(map(float, p.dims))} m room, full-ray activation",
            "updates_per_step": n_upd,
            "window_updates_per_step": n_win,
            "static_front_updates_per_step": n_static,
            "front_slots": st["front_slots"],
            "window_slots": st["window_slots"],
            "pairs_per_step": st["pairs_per_iter"],
            "rays": st["rays"],
            "list_records": st["list_records"],
            "ray_schedule_records_ahead_front_x_window": list(_lib.RAY_SCHEDULES[st["ray_sched"]]),
            "autotune_ms": st["t_autotune_ms"],
            "launch": a.launch,
            f"ms_per_step_{other}": dt_other * 1e3 / a.steps if dt_other is not None else None,
        },
        "time_to_converged": ttc,
        "reference_run": {"iters": conv["iters"], "residual": conv["residual"], "tol": a.tol,
                          "ms": conv["t_total_ms"]},
        "geometry_ms": conv["t_geometry_ms"],
        "roofline": {
            "kernel": "k_front_sums + k_ray_window (one camera pass)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": avg_launch_ms,
            "duration_source": "HIP events on the library stream around each camera's ray kernels, "
                               f"{'timed steps' if a.launch == 'timed' else 'comparison steps'}",
            "alg_bytes_per_launch": bytes_per_launch,
            "alg_bytes_per_front_update": front_b,
            "alg_bytes_per_window_update": window_b,
        },
        "cpu_baseline": cpu,
        "cpu_comparator": comp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
