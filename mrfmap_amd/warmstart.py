"""Learned warm start of the BP: training data from the method's own
converged maps, training on the GPU, held-out evaluation.

    python -m mrfmap_amd.warmstart train --out FILE [--small N] [--real N] [--batched N]
    python -m mrfmap_amd.warmstart eval [--weights FILE] [--workload small] [--seeds 2000-2019]
    python -m mrfmap_amd.warmstart convergence [--weights FILE] [--workload real] [--iters 5,10,...]

A training sample is one touched voxel of a synthetic map: its 32 features
after ``warm_after`` plain BP iterations (mrf_features) and, as target, the logit of its
marginal in the same map converged with plain FP32 BP (residual below 3e-5
or 400 iterations).  Training seeds are 1000-1199 of the workload generators of
mrfmap_amd.synth, held-out seeds 2000-2019.  Training and inference run
on the GPU (mrf_mlp_train / mrf_perform_inference_warm).

Voxels far from every surface are almost all saturated free space with the
same target; they are subsampled (``keep_free``) so that the voxels near
surfaces, where BP needs its iterations, carry the loss.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib, synth
from .inference import GVDBInference
from .params import gvdb_params

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warm_mlp.npz")
LF = 9.2102404
TRAIN_SEED0 = 1000
HELDOUT_SEEDS = tuple(range(2000, 2020))


def build(p, frames):
    g = GVDBInference(params=p)
    for pose, K, depth in frames:
        g.add_camera_with_depth(pose, depth, K)
    return g


def reference_map(p, frames, tol=3e-5, max_iters=400):
    """Marginals of plain FP32 BP iterated until the residual is below tol
    (FP32 Gauss-Seidel BP settles at residuals of 1e-5..1e-4 on the room
    workloads; some maps keep a small oscillation and stop at max_iters)."""
    g = build(p, frames)
    try:
        st = g.perform_inference(max_iters=max_iters, tol=tol)
        return g.get_alphas(), st
    finally:
        g.close()


def logit(a):
    a = np.clip(np.asarray(a, dtype=np.float64), 1e-4, 1 - 1e-4)
    return np.log(a) - np.log1p(-a)


def samples(p, frames, rng, keep_free=0.02, cap=None, ref=None, warm_after=1):
    """(X, t) of one map: features after warm_after iterations, logit of the converged marginal."""
    if ref is None:
        ref, _ = reference_map(p, frames)
    g = build(p, frames)
    try:
        g.perform_inference(max_iters=warm_after)
        X, vox = g.features()
        if g.num_bricks() != ref.shape[0]:
            raise RuntimeError("brick pools of the two runs differ")
    finally:
        g.close()
    t = logit(ref.ravel()[vox]).astype(np.float32)
    near = (t > -LF + 1e-3) | (X[:, 0] > 0) | (X[:, 4] < 1.0)
    idx = np.nonzero(near | (rng.random(len(t)) < keep_free))[0]
    if cap and len(idx) > cap:
        idx = np.sort(rng.choice(idx, cap, replace=False))
    return X[idx], t[idx]


def collect(spec, seed0=TRAIN_SEED0, keep_free=0.02, caps=None, log=print, warm_after=1):
    """spec: {workload name: number of seeds}; seeds seed0, seed0 + 1, ..."""
    caps = caps or {}
    rng = np.random.default_rng(12345)
    Xs, ts = [], []
    for name, nseeds in spec.items():
        for k in range(nseeds):
            seed = seed0 + k
            t0 = time.perf_counter()
            kw = {"frames": 8} if name == "batched" else {}
            p, frames, _ = synth.make_workload(name, seed=seed, **kw)
            X, t = samples(p, frames, rng, keep_free=keep_free, cap=caps.get(name), warm_after=warm_after)
            Xs.append(X)
            ts.append(t)
            log(f"[warmstart] {name} seed {seed}: {len(t)} samples ({time.perf_counter() - t0:.1f} s)")
    return np.concatenate(Xs), np.concatenate(ts)


def device_context():
    """A context without keyframes, for the MLP calls."""
    p = gvdb_params()
    p.rows, p.cols = 8, 8
    p.K = np.array([[4, 0, 4], [0, 4, 4], [0, 0, 1]], dtype=np.float32)
    return GVDBInference(params=p)


def init_params(seed=0):
    """He-uniform weights, zero biases, output bias at the free-space logit."""
    rng = np.random.default_rng(seed)
    p = np.zeros(_lib.MLP_PARAMS, np.float32)
    i, h = _lib.MLP_FEATURES, 64
    w1, w2 = h * i, h * h
    p[:w1] = rng.uniform(-1, 1, w1) * np.sqrt(6.0 / i)
    p[w1 + h:w1 + h + w2] = rng.uniform(-1, 1, w2) * np.sqrt(6.0 / h)
    p[w1 + h + w2 + h:w1 + h + w2 + 2 * h] = rng.uniform(-1, 1, h) * np.sqrt(6.0 / h)
    p[-1] = -LF
    return p


def train(X, t, epochs=12, batch=8192, lr=2e-3, lr_decay=0.8, seed=0, log=print):
    g = device_context()
    try:
        p0 = init_params(seed)
        p, losses = g.mlp_train(p0, X, t, epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay, seed=seed + 1)
    finally:
        g.close()
    for e, l in enumerate(losses):
        log(f"[warmstart] epoch {e}: mean loss {l:.4f}")
    return p, losses


def save_weights(path, params, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), meta=np.array(json.dumps(meta)))


def load_weights(path=None):
    """(params float32[6337], metadata dict) of a trained network; the
    metadata's ``warm_after`` is the iteration count its features are taken at."""
    with np.load(path or DEFAULT_WEIGHTS, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if params.shape != (_lib.MLP_PARAMS,):
        raise ValueError(f"{path}: expected {_lib.MLP_PARAMS} parameters")
    return params, meta


def time_to_error(p, frames, ref, thresh=1e-3, params=None, max_iters=500, growth=2.0, warm_after=1):
    """Wall time from keyframes added to a map within ``thresh`` max-abs of
    ``ref``: geometry, iterations (and, with ``params``, features, MLP and
    warm-start update).  The error check runs on the GPU after every
    iteration and is included."""
    g = build(p, frames)
    try:
        t0 = time.perf_counter()
        g.prepare()
        t_geom = (time.perf_counter() - t0) * 1e3
        g.set_reference(ref, stop_at=thresh)
        if params is None:
            st = g.perform_inference(max_iters=max_iters)
        else:
            st = g.perform_inference_warm(params, max_iters=max_iters, growth=growth, warm_after=warm_after)
        resid, err = g.get_trace()
    finally:
        g.close()
    return {"ms": t_geom + st["t_total_ms"], "geometry_ms": t_geom, "iters": st["iters"],
            "error": st["ref_error"], "reached": bool(st["converged"]), "warm": st["warm"],
            "warm_ms": st["t_warm_ms"], "autotune_ms": st["t_autotune_ms"],
            "trace_resid": [float(x) for x in resid], "trace_err": [float(x) for x in err]}


def evaluate(params, workload="small", seeds=HELDOUT_SEEDS, thresh=1e-3, msg=_lib.MSG_F32, log=print, warm_after=1,
             growth=2.0, **kw):
    rows = []
    for seed in seeds:
        p, frames, _ = synth.make_workload(workload, seed=seed, **kw)
        ref, _ = reference_map(p, frames)
        plain = time_to_error(p, frames, ref, thresh)
        p.msg_dtype = msg
        warm = time_to_error(p, frames, ref, thresh, params=params, warm_after=warm_after, growth=growth)
        rows.append({"seed": seed, "plain": plain, "warm": warm})
        log(f"[warmstart] {workload} seed {seed}: plain {plain['iters']} it {plain['ms']:.1f} ms, "
            f"warm {warm['iters']} it {warm['ms']:.1f} ms (state {warm['warm']}, error {warm['error']:.1e})")
    return rows


def convergence(p, frames, ref, iters, params=None, warm_after=1, thresh=1e-3):
    """Per voxel, the first of the sampled iteration counts ``iters`` from
    which on its marginal stays within ``thresh`` of ``ref`` (a fresh context
    per count: a call resets the camera messages, so k iterations of one call
    are the state after k).  Returns (index into iters per voxel, len(iters)
    for a voxel still outside at the last count; the max error per count)."""
    E = []
    for k in iters:
        g = build(p, frames)
        try:
            if params is None:
                g.perform_inference(max_iters=k)
            else:
                g.perform_inference_warm(params, max_iters=k, warm_after=warm_after, growth=0.0)
            E.append(np.abs(g.get_alphas().ravel() - ref.ravel()))
        finally:
            g.close()
    bad = np.stack(E) > thresh
    # some[k]: outside at count k or later; the first count from which on a
    # voxel stays inside is the number of such k
    some = np.cumsum(bad[::-1], axis=0)[::-1] > 0
    return some.sum(axis=0), [float(e.max()) for e in E]


def convergence_report(workload="real", seed=None, iters=(5, 10, 20, 40, 60, 80, 100, 120, 140, 160, 200),
                       params=None, warm_after=1, thresh=1e-3, log=print):
    """Where the iterations to a map within thresh of the converged one go:
    histograms of the per-voxel convergence count for plain and (with
    params) warm-started BP, and, for the voxels that converge last, their
    converged marginal and their distance to the nearest depth endpoint."""
    kw = {} if seed is None else {"seed": seed}
    p, frames, _ = synth.make_workload(workload, **kw)
    ref, st = reference_map(p, frames)
    log(f"[warmstart] reference: {st['iters']} iterations, residual {st['residual']:.2e}")
    g = build(p, frames)
    try:
        g.perform_inference(max_iters=1)
        X, vox = g.features()
    finally:
        g.close()
    dist = np.full(ref.size, -1.0, np.float32)
    dist[vox] = X[:, 4] * 4.0  # distance to the nearest endpoint voxel, capped at 4 voxels
    out = {"workload": workload, "thresh": thresh, "iters": list(iters), "ref_iters": int(st["iters"]),
           "ref_residual": float(st["residual"])}
    runs = [("plain", None)] + ([("warm", params)] if params is not None else [])
    for name, prm in runs:
        k, emax = convergence(p, frames, ref, iters, params=prm, warm_after=warm_after, thresh=thresh)
        hist = np.bincount(k, minlength=len(iters) + 1)
        late = k >= np.searchsorted(iters, 40)  # converging after 40 iterations or not at all
        rl = ref.ravel()[late]
        dl = dist[late]
        rep = {"hist": [int(x) for x in hist], "max_error": emax, "late_voxels": int(late.sum()),
               "late_ref_marginal_quantiles": [float(x) for x in np.quantile(rl, [0.1, 0.5, 0.9])] if late.any() else [],
               "late_by_distance": {str(d): int((dl == d).sum()) for d in (0.0, 1.0, 2.0, 3.0, 4.0)},
               "late_untouched_by_features": int((dl < 0).sum())}
        out[name] = rep
        log(f"[warmstart] {name}: converged-by histogram {rep['hist']} (last bin: not within {thresh} at "
            f"{iters[-1]}); max errors {['%.1e' % e for e in emax]}")
    return out


def _seeds(s):
    a, _, b = s.partition("-")
    return tuple(range(int(a), int(b) + 1)) if b else (int(a),)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--small", type=int, default=200)
    t.add_argument("--real", type=int, default=20)
    t.add_argument("--batched", type=int, default=0)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--keep-free", type=float, default=0.02)
    t.add_argument("--cap-small", type=int, default=60000)
    t.add_argument("--cap-real", type=int, default=600000)
    t.add_argument("--eval-seeds", default="2000-2004")
    t.add_argument("--warm-after", type=int, default=1)
    cv = sub.add_parser("convergence")
    cv.add_argument("--weights", default=None)
    cv.add_argument("--workload", default="real")
    cv.add_argument("--iters", default="5,10,20,40,60,80,100,120,140,160,200")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=None)
    e.add_argument("--workload", default="small")
    e.add_argument("--seeds", default="2000-2019")
    e.add_argument("--msg", type=int, default=_lib.MSG_F32)
    e.add_argument("--growth", type=float, default=2.0)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.cmd == "train":
        t0 = time.perf_counter()
        spec = {k: v for k, v in (("small", a.small), ("real", a.real), ("batched", a.batched)) if v > 0}
        X, tt = collect(spec, keep_free=a.keep_free,
                        caps={"small": a.cap_small, "real": a.cap_real, "batched": a.cap_real}, log=log,
                        warm_after=a.warm_after)
        t_data = time.perf_counter() - t0
        log(f"[warmstart] {len(tt)} samples in {t_data:.1f} s")
        t1 = time.perf_counter()
        params, losses = train(X, tt, epochs=a.epochs, log=log)
        t_train = time.perf_counter() - t1
        meta = {"features": _lib.MLP_FEATURES, "layers": [32, 64, 64, 1], "samples": int(len(tt)),
                "seeds": {k: [TRAIN_SEED0, TRAIN_SEED0 + v - 1] for k, v in spec.items()},
                "keep_free": a.keep_free, "epochs": a.epochs, "warm_after": a.warm_after, "losses": [float(x) for x in losses],
                "data_s": t_data, "train_s": t_train}
        save_weights(a.out, params, meta)
        log(f"[warmstart] trained in {t_train:.1f} s -> {a.out}")
        rows = evaluate(params, "small", _seeds(a.eval_seeds), log=log, warm_after=a.warm_after)
        print(json.dumps({"eval_small": rows}))
        return 0
    if a.cmd == "convergence":
        params, meta = load_weights(a.weights)
        rep = convergence_report(a.workload, iters=tuple(int(x) for x in a.iters.split(",")), params=params,
                                 warm_after=meta.get("warm_after", 1), log=log)
        print(json.dumps(rep))
        return 0
    params, meta = load_weights(a.weights)
    rows = evaluate(params, a.workload, _seeds(a.seeds), msg=a.msg, log=log, warm_after=meta.get("warm_after", 1),
                    growth=a.growth)
    pi = np.array([r["plain"]["iters"] for r in rows])
    wi = np.array([r["warm"]["iters"] for r in rows])
    pm = np.array([r["plain"]["ms"] for r in rows])
    wm = np.array([r["warm"]["ms"] for r in rows])
    print(json.dumps({"workload": a.workload, "seeds": a.seeds, "plain_iters_mean": pi.mean(),
                      "warm_iters_mean": wi.mean(), "plain_ms_mean": pm.mean(), "warm_ms_mean": wm.mean(),
                      "speedup_mean": float((pm / wm).mean()), "fallbacks": int(sum(r["warm"]["warm"] == 2 for r in rows)),
                      "rows": rows}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
