"""Writes tests/golden/c2_anchor.npz: the FP64 oracle's converged map of the
640x480 / 1 cm / 13-keyframe workload with every STEP-th pixel row and
column kept (the others invalid), the anchor against which
tests/test_gpu_parity.py checks the GPU's converged FP32 map.

The oracle (FP64, all host cores) iterates plain BP from the prior until the
max marginal change of an iteration is below TOL (or MAX_ITERS).  The map is
stored compactly: the pool indices of the voxels whose marginal exceeds
THRESH (sorted, as uint32 deltas) with their marginals rounded to 1/65535,
and the largest marginal of all other voxels, which bounds them.  The brick
coordinates of the pool (in slot order) are stored too, so that a reader can
check that its brick pool is the same.

    python tests/golden/make_c2_anchor.py [STEP] [THRESH]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

STEP = int(sys.argv[1]) if len(sys.argv) > 1 else 3
THRESH = float(sys.argv[2]) if len(sys.argv) > 2 else 0.02
TOL = 1e-6
MAX_ITERS = 400


def main():
    import oracle
    from helpers import c2_anchor_workload, oracle_for

    p, frames = c2_anchor_workload(STEP)
    o = oracle_for(p, frames)
    t0 = time.time()
    o.prepare()
    print(f"prepare {time.time() - t0:.0f} s, {o.num_bricks()} bricks, {oracle.get_threads()} threads", flush=True)
    st = o.perform_inference(10, tol=TOL)
    iters = st["iters"]
    while st["residual"] >= TOL and iters < MAX_ITERS:
        st = o.continue_inference(min(20, MAX_ITERS - iters), tol=TOL)
        iters += st["iters"]
        print(f"{iters} iterations, residual {st['residual']:.3e}, {time.time() - t0:.0f} s", flush=True)
    a = o.alphas().ravel()
    keep = np.nonzero(a > THRESH)[0].astype(np.uint32)
    deltas = np.diff(keep, prepend=np.uint32(0)).astype(np.uint32)
    vals = np.round(a[keep] * 65535.0).astype(np.uint16)
    rest = np.ones(a.size, bool)
    rest[keep] = False
    meta = {"workload": "real", "frames": 13, "step": STEP, "brick_local_vis": int(p.brick_local_vis),
            "iters": int(iters), "residual": float(st["residual"]), "tol": TOL, "thresh": THRESH,
            "num_bricks": int(o.num_bricks()), "kept": int(len(keep)), "rest_max": float(a[rest].max()),
            "threads": oracle.get_threads()}
    out = os.path.join(HERE, "c2_anchor.npz")
    np.savez_compressed(out, deltas=deltas, vals=vals, bricks=o.brick_coords().astype(np.int16),
                        meta=np.array(json.dumps(meta)))
    print(json.dumps(meta), f"-> {out} ({os.path.getsize(out) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
