"""Shared helpers of the test-suite (inputs, oracle runs, comparisons)."""
import numpy as np

import oracle
from mrfmap_amd import GVDBInference, synth


def decimate(depth, step):
    """Keep every `step`-th row and column of a depth image (others invalid)."""
    out = np.zeros_like(depth)
    out[::step, ::step] = depth[::step, ::step]
    return out


def oracle_for(p, frames, activation=None):
    tables = {}
    if p.use_polys and p.depth_sigma_lookup is not None and p.depth_bias_lookup is not None:
        tables = {"sigma_lookup": p.depth_sigma_lookup, "bias_lookup": p.depth_bias_lookup}
    o = oracle.Oracle(p.K, p.rows, p.cols, p.res, p.prior, p.dims, p.const_sigmasq_poly, p.const_bias_poly,
                      lookup_n=p.lookup_n, activation=p.activation if activation is None else activation,
                      mask_848_border=int(p.mask_848_border),
                      brick_local_vis=int(p.brick_local_vis), **tables)
    for pose, K, depth in frames:
        o.add_camera(pose, depth, K)
    return o


def gpu_for(p, frames):
    g = GVDBInference(params=p)
    for pose, K, depth in frames:
        g.add_camera_with_depth(pose, depth, K)
    return g


def small(**kw):
    return synth.make_workload("small", **kw)


def coffin_frames(activation, brick_local_vis=0, root=None):
    """The reference's own inputs: keyframes 0 and 6 of data/ (640x480 planar
    depth, NaN in the top rows) with the parameters of
    config/coffin_world_640.yaml (res 0.02, K 381.36, sigma 0.05, prior 0.1,
    dims 4x4x3, max_iters 3), each pose composed as pose * cam_in_body
    (testMRFMapViewer.cpp:55)."""
    import os

    from mrfmap_amd import gvdb_params

    gold = os.path.join(root or os.path.dirname(os.path.abspath(__file__)), "golden")
    p = gvdb_params().load_from_file(os.path.join(gold, "coffin_world_640.yaml"))
    p.activation = activation
    p.brick_local_vis = brick_local_vis
    z = np.load(os.path.join(gold, "coffin_frames.npz"))
    frames = [((z["body_pose"][i] @ z["cam_in_body"]).astype(np.float32), None, z["depth"][i])
              for i in range(len(z["frames"]))]
    return p, frames


def c2_anchor_workload(step):
    """The keyframes of the FP64 anchor (tests/golden/c2_anchor.npz): the
    640x480 / 1 cm / 13-keyframe workload with every step-th pixel row and
    column kept."""
    p, frames, _ = synth.make_workload("real", frames=13)
    return p, [(pose, K, decimate(d, step)) for pose, K, d in frames]
