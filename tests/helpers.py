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
                      mask_848_border=int(p.mask_848_border), **tables)
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
