"""Pins the oracle's map evaluation (orc_evaluate) with closed forms.

A 16^3 grid of 10 cm voxels, a camera on the world z axis at z = 0.01 m
looking along +z, a 5 x 5 image whose centre pixel's ray runs straight up
the voxel column (8, 8, *): chord length 1 voxel per voxel (0.9 for the
first), exit depth (k + 1) 0.1 - 0.01 m for voxel k.

* unknown map (no voxel updated): visibility stays 1, so every valid pixel
  has log-likelihood log(LIKE_EPS), accuracy -1 (no weight peak), expected
  depth 0, and simple accuracy 0 (an active brick ends beyond the
  measurement before any voxel reaches acc_thresh) — gvdbkernels.cu:135-181,
  1066-1104, 1246-1274, 1332-1363;
* one occupied voxel (marginal 1 - F_EPS, alpha = -ln F_EPS, so the
  visibility behind it is F_EPS): expected depth = (1 - F_EPS) x its centre
  depth; likelihood = (1 - F_EPS) (1 - LIKE_EPS) (N(d; s, 3 sigma) +
  N(d; s_prev, 3 sigma)) / 2 + F_EPS LIKE_EPS with the normal pdf; accuracy 1
  and simple accuracy 1 (true positive) when it is within sigma / one voxel
  of the measurement, and 2 (false positive) for the simple test when it is
  not.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle

RES = 0.1
SIG = 0.03
RES32 = float(np.float32(RES))  # the map's voxel size and the camera centre enter as float32
TZ32 = float(np.float32(0.01))
F_EPS = 1e-4
LIKE_EPS = 1e-3
D = 0.8


def make(depth=D):
    K = np.array([[10, 0, 2], [0, 10, 2], [0, 0, 1]], dtype=np.float32)
    o = oracle.Oracle(K, 5, 5, RES, 0.1, [1.6, 1.6, 1.6], [SIG, 0, 0], [0, 0, 0], lookup_n=3, activation=1)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [0.05, 0.05, 0.01]
    dep = np.full((5, 5), depth, dtype=np.float32)
    o.add_camera(pose, dep, K)
    o.prepare()
    return o, pose, dep, K


def axis_voxel(o, k):
    """Pool index of voxel (8, 8, k) and the brick slot table order."""
    coords = o.brick_coords()  # slot -> brick (bx, by, bz)
    slot = [i for i, b in enumerate(coords) if tuple(b) == (1, 1, k // 8)]
    assert len(slot) == 1
    return slot[0] * 512 + ((8 & 7) * 8 + (8 & 7)) * 8 + (k & 7)


def test_unknown_map():
    o, pose, dep, K = make()
    n = o.num_bricks()
    o.set_beliefs(np.full((n, 512), math.log(0.1 / 0.9)), np.zeros((n, 512), np.uint8))
    r = o.evaluate(pose, dep, K)
    np.testing.assert_allclose(r["loglik"], math.log(LIKE_EPS), rtol=0, atol=1e-12)
    assert (r["accuracy_image"] == -1).all()
    assert (r["expected_depth"] == 0).all()
    assert (r["simple_accuracy_image"] == 0).all()
    assert r["accuracy"] == 0.0 and r["simple_counts"] == [0, 25, 0, 0]
    assert math.isclose(r["log_likelihood"], 25 * math.log(LIKE_EPS), rel_tol=1e-12)


@pytest.mark.parametrize("k", [7, 6, 3])
def test_one_occupied_voxel(k):
    o, pose, dep, K = make()
    n = o.num_bricks()
    B = np.full((n, 512), math.log(0.1 / 0.9))
    upd = np.zeros((n, 512), np.uint8)
    v = axis_voxel(o, k)
    B.flat[v] = 30.0  # marginal clamps to 1 - F_EPS
    upd.flat[v] = 1
    o.set_beliefs(B, upd)
    r = o.evaluate(pose, dep, K)
    s, s_prev, zc = (k + 1) * RES32 - TZ32, k * RES32 - TZ32, (k + 0.5) * RES32 - TZ32
    assert math.isclose(r["expected_depth"][2, 2], (1 - F_EPS) * zc, rel_tol=1e-9)
    if abs(s - D) <= 3 * SIG:
        pz = (1 - LIKE_EPS) * 0.5 * (stats.norm.pdf(D, s, 3 * SIG) + stats.norm.pdf(D, s_prev, 3 * SIG))
    else:
        pz = LIKE_EPS
    want = math.log((1 - F_EPS) * pz + F_EPS * LIKE_EPS)
    assert math.isclose(r["loglik"][2, 2], want, rel_tol=1e-5)
    assert r["accuracy_image"][2, 2] == (1 if abs(s - D) <= SIG else 0)
    assert r["simple_accuracy_image"][2, 2] == (1 if abs(s - D) <= RES32 else 2)


def test_evaluation_of_a_converged_map_prefers_the_true_pose():
    """Sanity of the full path on the synthetic room (4 keyframes, 10 BP
    iterations): a keyframe's depth image is more likely at its own pose
    than 12 cm away, the class counts cover the image, and the expected
    depth follows the measurement."""
    from mrfmap_amd import synth
    from helpers import oracle_for

    p, frames, _ = synth.make_workload("small", frames=4)
    o = oracle_for(p, frames)
    o.perform_inference(10)
    pose, K, depth = frames[0]
    r0 = o.evaluate(pose, depth, K, acc_thresh=p.acc_thresh)
    moved = pose.copy()
    moved[0, 3] += 0.12
    r1 = o.evaluate(moved, depth, K, acc_thresh=p.acc_thresh)
    assert r0["log_likelihood"] > r1["log_likelihood"] + 1000
    assert sum(r0["accuracy_counts"]) == depth.size == sum(r0["simple_counts"])
    assert r0["accuracy"] > 0.3
    valid = depth > 0
    err = np.abs(r0["expected_depth"][valid] - depth[valid])
    assert np.median(err) < 0.2


def test_load_map_single_point_closed_form():
    """load_map (GVDBMapLikelihoodEstimator.cpp:168-240): one point of
    occupancy 0.9999 in voxel (8, 8, 7) on the axis ray gives the closed-form
    expected depth of an occupied voxel; a point outside the map box is
    ignored; bricks within one voxel of the point become active."""
    K = np.array([[10, 0, 2], [0, 10, 2], [0, 0, 1]], dtype=np.float32)
    o = oracle.Oracle(K, 5, 5, RES, 0.1, [1.6, 1.6, 1.6], [SIG, 0, 0], [0, 0, 0], lookup_n=3, activation=1)
    o.load_map(np.array([[8.5, 8.5, 7.5], [-3.0, 2.0, 2.0]], np.float32), np.array([0.9999, 0.7], np.float32))
    # bricks (1, 1, 0) plus the neighbours of voxel 8.5 +- 1 -> x, y bricks 0..1, z brick 0..1
    assert o.num_bricks() == 8
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [0.05, 0.05, 0.01]
    dep = np.full((5, 5), D, dtype=np.float32)
    r = o.evaluate(pose, dep, K)
    zc = 7.5 * RES32 - TZ32
    vis = 1.0 - float(np.float32(0.9999))
    assert math.isclose(r["expected_depth"][2, 2], (1 - vis) * zc / (1 - vis + F_EPS), rel_tol=1e-6)
