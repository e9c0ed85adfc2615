"""Pins the oracle's warm-start features (orc_features) and warm-start
update (orc_warm_start) on a one-camera scene.

The scene: 16^3 grid of 10 cm voxels, a 5 x 5 camera (f = 10, c = (2, 2)) at
(0.05, 0.05, 0.01) m looking along +z, depth 0.8 m everywhere, one BP
iteration from the prior.  Every pixel's endpoint lands in its own voxel
(x, y) = (floor(8.5 + 0.8 (u - 2)), floor(8.5 + 0.8 (v - 2))), z = 8, found
here from the camera geometry directly.  Checks:

* one camera: the message mean equals the message max, and the belief is
  the prior plus that message (B = L0 + M_c after one iteration);
* endpoint counts against the 25 endpoint voxels;
* the distance feature against a brute-force search over those voxels;
* the chord-length feature against the sum of the camera's traversal
  records of the voxel;
* the 27 neighbourhood features against the beliefs looked up by brick
  coordinates;
* orc_warm_start: beliefs move to the prediction, B = L0 + sum_c M_c still
  holds (next iteration from the moved state equals a fresh evaluation).
"""
import math

import numpy as np

import oracle

RES = 0.1
LF = 9.2102404


def make():
    K = np.array([[10, 0, 2], [0, 10, 2], [0, 0, 1]], dtype=np.float32)
    o = oracle.Oracle(K, 5, 5, RES, 0.1, [1.6, 1.6, 1.6], [0.03, 0, 0], [0, 0, 0], lookup_n=3, activation=1)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [0.05, 0.05, 0.01]
    o.add_camera(pose, np.full((5, 5), 0.8, dtype=np.float32), K)
    o.prepare()
    return o


def coords_of(o, vox):
    bc = o.brick_coords()
    s, t = vox // 512, vox % 512
    return np.stack([bc[s, 0] * 8 + (t >> 6), bc[s, 1] * 8 + ((t >> 3) & 7), bc[s, 2] * 8 + (t & 7)], axis=1)


def endpoint_voxels():
    return {(int(math.floor(8.5 + 0.8 * (u - 2))), int(math.floor(8.5 + 0.8 * (v - 2))), 8)
            for u in range(5) for v in range(5)}


def test_features_one_camera():
    o = make()
    o.perform_inference(1)
    X, vox = o.features()
    assert len(X) > 50
    xyz = coords_of(o, vox.astype(np.int64))
    B = o.beliefs().ravel()
    L0 = math.log(0.1 / 0.9)
    # one camera: mean = max, B = L0 + M
    np.testing.assert_array_equal(X[:, 2], X[:, 3])
    np.testing.assert_allclose(X[:, 2], ((B[vox] - L0) / LF).astype(np.float32), rtol=1e-6, atol=1e-7)
    # endpoint counts
    hits = endpoint_voxels()
    assert len(hits) == 25
    for row, c in enumerate(map(tuple, xyz)):
        assert X[row, 0] == np.float32(math.log1p(1.0 if c in hits else 0.0)), c
    assert sum(1 for c in map(tuple, xyz) if c in hits) == 25
    # distance to the nearest endpoint voxel, capped at 4 voxels
    H = np.array(sorted(hits))
    for row in range(len(X)):
        d = np.sqrt(((H - xyz[row]) ** 2).sum(1)).min()
        assert abs(X[row, 4] - min(d, 4.0) / 4.0) < 1e-6, (xyz[row], d, X[row, 4])
    # chord length from the traversal records
    _, recs = o.traversal(0)
    live = recs[:, 0] != 0xFFFFFFFF
    L = np.zeros(o.num_bricks() * 512)
    np.add.at(L, recs[live, 0].astype(np.int64), recs[live, 1].view(np.float32).astype(np.float64))
    np.testing.assert_allclose(X[:, 1], np.log1p(L[vox]).astype(np.float32), rtol=1e-6)
    # neighbourhood beliefs by coordinates
    allc = coords_of(o, np.arange(o.num_bricks() * 512))
    lut = {tuple(c): B[i] for i, c in enumerate(allc)}
    f = 5
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for e in (-1, 0, 1):
                want = np.array([lut.get((x + a, y + b, z + e), L0) for x, y, z in xyz]) / LF
                np.testing.assert_allclose(X[:, f], np.clip(want, -3, 3).astype(np.float32), rtol=1e-6, atol=1e-7)
                f += 1
    assert f == 32


def test_warm_start_moves_beliefs_and_keeps_message_sum():
    """After orc_warm_start(pred) the beliefs are the predictions, and one
    more iteration gives the same beliefs as starting a context whose state
    was set to those beliefs with the camera messages moved by the same
    amount: with one camera, M_c = B - L0 exactly, so the next message only
    depends on the prediction through q = B - M_c = L0 (the cavity of a
    single camera is the prior), i.e. the continuation equals plain BP's
    second iteration."""
    o = make()
    o.perform_inference(1)
    X, vox = o.features()
    pred = np.linspace(-3, 3, len(vox))
    o.warm_start(vox, pred)
    B = o.beliefs().ravel()
    np.testing.assert_allclose(B[vox], pred, rtol=0, atol=1e-12)
    o.continue_inference(1)
    ref = make()
    ref.perform_inference(2)
    np.testing.assert_allclose(o.beliefs(), ref.beliefs(), rtol=0, atol=1e-9)


def test_warm_start_keeps_stronger_saturated_belief():
    o = make()
    o.perform_inference(1)
    X, vox = o.features()
    B0 = o.beliefs().ravel()[vox]
    free = np.where(B0 < -LF - 1)[0]
    assert len(free) > 0
    pred = np.full(len(vox), -LF - 0.5)  # saturated free, weaker than the first iteration there
    o.warm_start(vox, pred)
    B = o.beliefs().ravel()[vox]
    np.testing.assert_array_equal(B[free], B0[free])
    other = np.setdiff1d(np.arange(len(vox)), free[B0[free] <= -LF - 0.5])
    np.testing.assert_allclose(B[other], np.where(B0[other] <= pred[other], B0[other], pred[other]))
