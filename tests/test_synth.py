"""The synthetic keyframe generator: determinism, closed-form depth and the
noise model sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 with 2 % dropout."""
import numpy as np
import pytest

from mrfmap_amd import synth


def look_x_pose(x=0.0, y=0.0, z=1.5):
    """camera at (x, y, z) looking along +x (right = -y, down = -z)."""
    T = np.eye(4, dtype=np.float32)
    T[:3, 0] = [0, -1, 0]
    T[:3, 1] = [0, 0, -1]
    T[:3, 2] = [1, 0, 0]
    T[:3, 3] = [x, y, z]
    return T


def test_deterministic():
    a = synth.make_workload("small")[1]
    b = synth.make_workload("small")[1]
    for (p1, _, d1), (p2, _, d2) in zip(a, b):
        np.testing.assert_array_equal(p1, p2)
        np.testing.assert_array_equal(d1, d2)


def test_planar_depth_of_a_facing_wall():
    sc = synth.Scene("room", 5, num_boxes=0, num_spheres=0)
    K = synth.intrinsics(120, 160)
    d = sc.render(look_x_pose(-0.5), K, 120, 160, noise=False)
    # wall x = +2 seen head-on: planar depth 2.5 everywhere it is visible
    center = d[40:80, 50:110]
    np.testing.assert_allclose(center, 2.5, rtol=1e-6)


def test_ring_poses_look_at_centre():
    sc = synth.Scene("room", 1)
    for k in range(4):
        T = sc.ring_pose(k, 4)
        assert np.hypot(T[0, 3], T[1, 3]) == pytest.approx(2.0, abs=1e-6)
        f = T[:3, 2]
        to_c = -T[:3, 3].copy()
        to_c[2] = 0
        to_c /= np.linalg.norm(to_c)
        assert np.degrees(np.arccos(np.clip(f @ to_c, -1, 1))) <= 5.0 + 1e-4
        R = T[:3, :3]
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-6)


def test_noise_model_and_dropout():
    sc = synth.Scene("room", 5, num_boxes=0, num_spheres=0)
    K = np.array([[20.0, 0, 7.5], [0, 20.0, 7.5], [0, 0, 1]])
    T = look_x_pose(-0.5)
    clean = sc.render(T, K, 16, 16, noise=False)
    stack = np.stack([sc.render(T, K, 16, 16, noise=True, dropout=0.02, noise_seed=s) for s in range(400)])
    drop = (stack == 0).mean()
    assert drop == pytest.approx(0.02, abs=0.004)
    valid = np.where(stack > 0, stack, np.nan)
    z = clean[8, 8]
    sig = 0.0012 + 0.0019 * (z - 0.4) ** 2
    resid = (valid - clean[None]) / sig
    assert np.nanmean(resid) == pytest.approx(0.0, abs=0.03)
    assert np.nanstd(resid) == pytest.approx(1.0, abs=0.03)


def test_workload_shapes():
    p, frames, _ = synth.make_workload("small")
    assert len(frames) == 4 and frames[0][2].shape == (120, 160)
    assert p.res == 0.04 and p.const_sigmasq_poly[0] == pytest.approx(0.06)
    d = frames[0][2]
    assert 0.01 < (d == 0).mean() < 0.04
