#!/usr/bin/env python3
"""Benchmark of the MRFMap BP inference path on one MI355X.

Workload "real": 13 synthetic keyframes of 640x480 (f=525, c=(319.5,239.5)),
1 cm voxels in a 4x4x3 m room with 12 boxes and 4 spheres (seed 2), full-ray
brick activation, FP32 messages, the reference's pre pass (visibility
restarted at brick entries until a tau survives flush-to-zero; --vis carry
for the carried visibility).  A step is one BP iteration: every camera's
ray-message pass + epilogue + the FP64 residual, with all keyframes, lists
and map state resident in HBM.

Prints one JSON line: updates/s over the timed steps (graph replay) and the
window-message update rate; per-kernel durations (HIP events on the
library's stream, steps launched one by one) with each kernel's algorithmic
bytes and roofline fraction; a 3-iteration call (the reference's default
max_iters) from its per-call message reset; the time to a residual below
1e-4; the time-to-converged map -- from "keyframes uploaded" (traversal
included) to a map within 1e-3 max-abs marginal error of plain FP32 BP
iterated to a residual below 3e-5 (the FP64 map of the decimated workload is
a committed fixture, tests/golden/c2_anchor.npz; FP32 and FP64 converged
maps agree to 4e-5 there) -- for plain FP32 BP, for the learned warm start
(mrfmap_amd/data/warm_mlp.npz) with FP32 messages and with FP16 ray->voxel
messages; the CPU baseline: the FP64 oracle on a bounded sample of the same
workload, on one host core and on all of them, with the host's CPU model.
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
    ap.add_argument("--vis", default="default", choices=["default", "carry", "brick-local"],
                    help="pre-pass visibility: the library's default, carried along the ray, or restarted "
                         "at every brick until a tau survives flush-to-zero (the reference's)")
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


def host_info():
    """CPU model and core counts of the host running the CPU legs."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(wl, vis):
    """FP64 scalar oracle on keyframes 0 and 6 of the same workload, every 4th
    pixel row; one BP iteration, timed on one host thread and on all the
    threads OpenMP gives it (OMP_NUM_THREADS, else every core)."""
    import oracle

    p, sub = _sample(wl)
    o = oracle.Oracle(p.K, p.rows, p.cols, p.res, p.prior, p.dims, p.const_sigmasq_poly, p.const_bias_poly,
                      lookup_n=p.lookup_n, activation=p.activation, brick_local_vis=vis)
    for pose, K, d in sub:
        o.add_camera(pose, d, K)
    o.prepare()
    rates = {}
    for nt in (1, 0):
        oracle.set_threads(nt)
        threads = oracle.get_threads()
        t0 = time.perf_counter()
        st = o.perform_inference(1)
        rates[threads if nt else "all"] = (st["updates_per_iter"] / (time.perf_counter() - t0), threads)
    oracle.set_threads(0)
    v1, _ = rates[1]
    va, nta = rates["all"]
    return {"value": va, "unit": "ray-msg updates/s", "cores": nta, "kind": "port", "value_1_core": v1,
            "host": host_info(),
            "sample": f"FP64 oracle BP, 1 iteration over keyframes 0 and 6 of '{wl}' with every 4th pixel row "
                      f"({st['updates_per_iter']} updates; {st['updates_per_iter'] / v1:.2f} s on 1 thread, "
                      f"{st['updates_per_iter'] / va:.2f} s on {nta})"}


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
    if a.vis != "default":
        p.brick_local_vis = _lib.VIS_CARRY if a.vis == "carry" else _lib.VIS_BRICK_LOCAL
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
    # the reference configuration's call: perform_inference with max_iters 3
    # (coffin_world_640.yaml:24), messages reset at the call; graph launches
    g.set_launch_mode(modes["graph"])
    calls3 = []
    for _ in range(3):
        t3 = time.perf_counter()
        s3 = g.perform_inference(max_iters=3)
        calls3.append((time.perf_counter() - t3) * 1e3)
    call3_ms = sorted(calls3)[1]
    g.close()  # one context at a time: the convergence runs below build their own
    if ref is not None:
        ttc = time_to_converged(p, frames, ref, a.legs.split(","))
    resid_leg = None
    if not a.no_converge:
        # the residual criterion of the 640x480 configuration: from keyframes
        # added (geometry included) until the max marginal change of an
        # iteration is below 1e-4 (FP64 reduction on the device)
        g = build()
        t0 = time.perf_counter()
        g.prepare()
        t_geom = (time.perf_counter() - t0) * 1e3
        sr = g.perform_inference(max_iters=400, tol=1e-4)
        g.close()
        resid_leg = {"tol": 1e-4, "iters": sr["iters"], "reached": bool(sr["converged"]),
                     "ms": t_geom + sr["t_total_ms"], "geometry_ms": t_geom}
        print(f"[bench] time to residual 1e-4: {sr['iters']} iters {resid_leg['ms']:.1f} ms", file=sys.stderr,
              flush=True)
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
    # ray kernels, algorithmic bytes: every incidence reads its 8 B record and
    # its pair's voxel->ray message; a window incidence also writes its
    # ray->voxel message (its 4 B segment position is read by k_seg_sum and
    # charged there).  Front records whose pair's message stops changing are
    # read in the first two iterations of a call only (afterwards their sums
    # are kept); the average over the K timed launches of a camera counts
    # them so.
    in_b = 4 if msg in (_lib.MSG_F32, _lib.MSG_MIXED_F16, _lib.MSG_MIXED_BF16) else 2
    out_b = 4 if msg == _lib.MSG_F32 else 2
    n_win = st["window_updates_per_iter"]
    n_static = st["static_front_updates"]
    front_b, window_b = 8 + in_b, 8 + in_b + out_b
    front_reads = (n_upd - n_win - n_static) + n_static * min(a.steps, 2) / a.steps
    bytes_per_launch = (front_reads * front_b + n_win * window_b) / ncam
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(a)
    # per kernel of a camera pass, each charged with the bytes it moves itself
    # (the same per-update figures; k_seg_sum reads each window position's
    # 4 B segment index and its message and writes one 4 B sum per window
    # pair (upper bound: one per pair); k_seg_update reads 28 B and writes
    # 8 B per window pair)
    npairs = st["pairs_per_iter"]
    per_kernel = {}
    kb = {
        "k_front_sums": (timed["t_front_ms"], front_reads * front_b),
        "k_ray_window": (timed["t_window_ms"], n_win * (8 + in_b + out_b)),
        "k_seg_sum": (timed["t_seg_sum_ms"], n_win * (4 + out_b) + 4.0 * min(npairs, n_win)),
        "k_seg_update": (timed["t_seg_update_ms"], None),
        "k_pair_msg": (timed["t_pair_msg_ms"], None),
    }
    for name, (tms, alg) in kb.items():
        ms = tms / launches
        ent = {"avg_launch_ms": ms}
        if alg is not None:
            gbs = alg / ncam / (ms * 1e-3) / 1e9
            ent.update({"alg_bytes_per_launch": alg / ncam, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS})
        per_kernel[name] = ent
    other_s = f"{dt_other * 1e3:.1f} ms" if dt_other is not None else "not run"
    print(f"[bench] {a.steps} steps ({a.launch}) in {dt * 1e3:.1f} ms, ({other}) {other_s}; "
          f"ray kernels {timed['t_ray_kernel_ms']:.1f} ms, epilogue {timed['t_epilogue_ms']:.1f} ms",
          file=sys.stderr, flush=True)

    cpu = comp = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a.workload, p.brick_local_vis)
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
            "pre_pass_visibility": "brick-local" if p.brick_local_vis == _lib.VIS_BRICK_LOCAL else "carry",
        },
        "window_updates_per_s": n_win * a.steps / dt,
        "static_front_share": n_static / n_upd if n_upd else 0.0,
        "call_3_iterations": {"ms": call3_ms, "updates_per_s": 3 * n_upd / (call3_ms * 1e-3),
                              "note": "perform_inference(max_iters=3) from its per-call message reset, graph "
                                      "launches, median of 3 calls: every pair's ln n is recomputed in the first "
                                      "two iterations of a call"},
        "time_to_residual": resid_leg,
        "time_to_converged": ttc,
        "reference_run": {"iters": conv["iters"], "residual": conv["residual"], "tol": a.tol,
                          "ms": conv["t_total_ms"]},
        "geometry_ms": conv["t_geometry_ms"],
        "geometry_note": "geometry_ms is the first build of the process: it also pays the HIP runtime's first use "
                         "and the first provisioning of the list memory; the build of a fresh context later in "
                         "the process is time_to_residual.geometry_ms",
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
            "alg_bytes_per_window_update": 8 + in_b + out_b,
            "per_kernel": per_kernel,
        },
        "cpu_baseline": cpu,
        "cpu_comparator": comp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
