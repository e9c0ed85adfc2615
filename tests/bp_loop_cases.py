"""Hand-computed cases of the BP loop, shared by the CPU test of the oracle
(test_oracle_bp_loop.py) and the GPU test of libmrfmap_hip
(test_gpu_parity.py).  The expected values are written out here from the
reference's formulas and do not go through oracle.c or the library:

* the camera message m = logit(clamp(S / C)) from the length-weighted mean
  probability of the camera's rays (gvdbkernels.cu:457-462, 950-951);
* the belief update B <- B + m - M_c, M_c <- m (gvdbkernels.cu:464-467);
* no message where the camera's summed chord length C <= 0.1, and a voxel
  never updated reads 0 (gvdbkernels.cu:408; GVDBMRFMap.cpp:154-208);
* the voxel->ray message excludes the camera's own previous message
  (gvdbkernels.cu:764-773);
* cameras are updated one after another inside an iteration, each reading
  the beliefs the previous one wrote (Gauss-Seidel, GVDBInference.cpp:155-167);
* every perform_inference call resets M_c and keeps B (GVDBInference.cpp:128-134).

The scenes are single rays along a grid axis, with 1 m voxels, so chord
lengths and exit depths are exact dyadic numbers.  A scene is built by a
`make(cams)` callable that returns an object with perform_inference(iters),
beliefs() and alphas() (flat, pool order) and num_bricks().
"""
import math

import numpy as np

F_EPS = 1e-4
PRIOR = 0.1
# float logit of the prior, as GVDBMRFMap.cpp:51 computes it
L0 = float(np.log(np.float32(PRIOR)) - np.log(np.float32(1 - PRIOR)))
# camera z -> world +x, camera x -> world +y, camera y -> world +z
R_X = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
SIG = 0.25  # sigma (m); the likelihood width is 3 sigma


def _clamp(x):
    return min(max(x, F_EPS), 1 - F_EPS)


def _logit(a):
    return math.log(a) - math.log(1 - a)


def _nin(b):
    """voxel->ray message P(o = 0) from the belief without this camera's message"""
    return _clamp(1.0 / (1.0 + math.exp(b)))


def _psi(s, d):
    u = (d - s) / (3 * SIG)
    return 0.398942 * math.exp(-0.5 * u * u) / (3 * SIG)


def _three_voxel_messages(n, l, s, d):
    """Reference-mode messages of a truncated 3-voxel ray, written out term by
    term: V_i = prod_{j<=i} n_j^l_j, tau_i = V_i (1 - n_i) psi_i, T = sum tau,
    mu1_i = P_i + V_i psi_i, mu0_i = P_i + R_i / n_i, p = clamp(mu1/(mu1+mu0+1e-7))."""
    V1 = n[0] ** l[0]
    V2 = V1 * n[1] ** l[1]
    V3 = V2 * n[2] ** l[2]
    p1, p2, p3 = _psi(s[0], d), _psi(s[1], d), _psi(s[2], d)
    t1, t2, t3 = V1 * (1 - n[0]) * p1, V2 * (1 - n[1]) * p2, V3 * (1 - n[2]) * p3
    mu1 = [V1 * p1, t1 + V2 * p2, t1 + t2 + V3 * p3]
    mu0 = [(t2 + t3) / n[0], t1 + t3 / n[1], t1 + t2 + 0.0 / n[2]]
    return [_clamp(mu1[i] / (mu1[i] + mu0[i] + 1e-7)) for i in range(3)]


def _pose(vx, vy, vz):
    """Camera-to-world pose (row-major) whose centre sits at voxel
    coordinates (vx, vy, vz) of the 16^3 grid below, looking along +x."""
    T = np.eye(4, dtype=np.float32)
    T[:3, :3] = R_X
    T[:3, 3] = [vx - 8.0, vy - 8.0, vz]
    return T


# The map: 16 x 16 x 16 m of 1 m voxels (map box x, y in [-8, 8), z in [0, 16)),
# keyframes of 1 x 1 pixel with K = identity, prior 0.1, sigma polynomial
# [SIG, 0, 0], no bias, full-ray activation.  A camera (x0, d) has its centre
# at voxel (x0, 8.5, 4.5) and depth d: its ray runs along +x to the end of the
# grid (x = 16).
K_UNIT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float32)
DIMS = [16.0, 16.0, 16.0]
RES = 1.0


def camera(x0, d):
    """(pose, depth image) of camera (x0, d)"""
    return _pose(x0, 8.5, 4.5), np.full((1, 1), d, dtype=np.float32)


def _voxel(o, x):
    """pool index of voxel (x, 8, 4): the only active brick is (1, 1, 0)"""
    assert o.num_bricks() == 1
    return (x - 8) * 64 + 0 * 8 + 4


# ray from x = 13.25: chords 0.75, 1, 1; exit depths 0.75, 1.75, 2.75.  The walk
# stops at the brick exit (2.75 >= d + 3 sigma = 2.625): no infinity term.
X0, L, S, D = 13.25, [0.75, 1.0, 1.0], [0.75, 1.75, 2.75], 1.875


def check_one_camera_one_iteration(make):
    """B = L0 + logit(clamp(p_i l_i / C_i)) with C_i = l_i, and the marginal
    1 - clamp(1/(1+e^B)).  Catches: S used instead of S/C (the first voxel's
    C is 0.75); a message or belief update of the wrong sign or form."""
    o = make([(X0, D)])
    o.perform_inference(1)
    B = np.asarray(o.beliefs(), dtype=np.float64).ravel()
    A = np.asarray(o.alphas(), dtype=np.float64).ravel()
    p = _three_voxel_messages([_nin(L0)] * 3, L, S, D)
    err = 0.0
    for i, x in enumerate((13, 14, 15)):
        want = L0 + _logit(_clamp(p[i] * L[i] / L[i]))
        err = max(err, abs(B[_voxel(o, x)] - want))
        err = max(err, abs(A[_voxel(o, x)] - (1 - _clamp(1 / (1 + math.exp(want))))))
    return err


def check_cum_length_gate(make):
    """A voxel whose summed chord length is C <= 0.1 gets no message: its
    belief stays L0 and it reads 0; it still attenuates the ray.  Ray from
    x = 13.9375: chords 0.0625, 1, 1.  Catches: the gate dropped or moved
    (C > 0 instead of C > 0.1)."""
    x0, d = 13.9375, 1.25  # exit depths 0.0625, 1.0625, 2.0625; stop: 2.0625 >= 2.0
    o = make([(x0, d)])
    o.perform_inference(1)
    B = np.asarray(o.beliefs(), dtype=np.float64).ravel()
    A = np.asarray(o.alphas(), dtype=np.float64).ravel()
    p = _three_voxel_messages([_nin(L0)] * 3, [0.0625, 1.0, 1.0], [0.0625, 1.0625, 2.0625], d)
    err = abs(B[_voxel(o, 13)] - L0) + abs(A[_voxel(o, 13)] - 0.0)
    for i, x in ((1, 14), (2, 15)):
        err = max(err, abs(B[_voxel(o, x)] - (L0 + _logit(p[i]))))
    return err


def check_gauss_seidel(make):
    """Two identical keyframes: the second camera's voxel->ray messages use
    the belief the first camera just wrote, B = L0 + m0 (its own M_c is
    still 0), so after one iteration B = L0 + m0 + m1 with m1 != m0.
    Catches: Jacobi order (every camera reading the beliefs from the start
    of the iteration gives B = L0 + 2 m0)."""
    o = make([(X0, D), (X0, D)])
    o.perform_inference(1)
    B = np.asarray(o.beliefs(), dtype=np.float64).ravel()
    p0 = _three_voxel_messages([_nin(L0)] * 3, L, S, D)
    m0 = [_logit(x) for x in p0]
    p1 = _three_voxel_messages([_nin(L0 + m) for m in m0], L, S, D)
    m1 = [_logit(x) for x in p1]
    assert min(abs(a - b) for a, b in zip(m0, m1)) > 0.05  # the two orders differ
    return max(abs(B[_voxel(o, x)] - (L0 + m0[i] + m1[i])) for i, x in enumerate((13, 14, 15)))


def check_fixed_point_of_one_camera(make):
    """One camera alone: its second iteration excludes its own message, sees
    the prior again and sends the same m, so B = L0 + m stays.  Catches: a
    missing -M_c in the belief update (B = L0 + 2m) and a voxel->ray message
    that does not exclude the camera's own message."""
    o = make([(X0, D)])
    o.perform_inference(2)
    B = np.asarray(o.beliefs(), dtype=np.float64).ravel()
    p = _three_voxel_messages([_nin(L0)] * 3, L, S, D)
    return max(abs(B[_voxel(o, x)] - (L0 + _logit(p[i]))) for i, x in enumerate((13, 14, 15)))


def check_second_call_resets_messages(make):
    """A second perform_inference call zeroes M_c but keeps B (Q4): its
    voxel->ray messages see B = L0 + m in full, and B = L0 + m + m'.
    Catches: M_c kept across calls (B would stay L0 + m) and B reset to the
    prior (B = L0 + m')."""
    o = make([(X0, D)])
    o.perform_inference(1)
    o.perform_inference(1)
    B = np.asarray(o.beliefs(), dtype=np.float64).ravel()
    m = [_logit(x) for x in _three_voxel_messages([_nin(L0)] * 3, L, S, D)]
    m2 = [_logit(x) for x in _three_voxel_messages([_nin(L0 + mi) for mi in m], L, S, D)]
    return max(abs(B[_voxel(o, x)] - (L0 + m[i] + m2[i])) for i, x in enumerate((13, 14, 15)))


CHECKS = {
    "one_camera_one_iteration": check_one_camera_one_iteration,
    "cum_length_gate": check_cum_length_gate,
    "gauss_seidel": check_gauss_seidel,
    "fixed_point_of_one_camera": check_fixed_point_of_one_camera,
    "second_call_resets_messages": check_second_call_resets_messages,
}
