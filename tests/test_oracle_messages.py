"""Pins the oracle's per-ray factor messages (orc_ray_messages).

* textbook mode vs brute-force enumeration of all 2^n occupancy
  configurations of the forward-sensor factor (n <= 11): this pins the
  O(n) bookkeeping of T, P_i, R_i, the mu0/mu1 decomposition and the
  infinity term;
* the depth likelihood psi is the reference's 0.398942 * N(d; s+bias, 3 sigma)
  (gvdbkernels.cu:57-72), checked against scipy's normal pdf and its integral;
* reference mode, hand-computed for n = 1 and n = 2, and its specific
  quirks: visibility includes the voxel itself and is length weighted
  (gvdbkernels.cu:781-783), the infinity term only when the walk ran out and
  d lies beyond the last voxel (gvdbkernels.cu:1196-1208), the 1e-7
  normaliser and the [F_EPS, 1-F_EPS] clamp (gvdbkernels.cu:942-943);
* invariants on a free-space ray with a surface.
"""
import math

import numpy as np
import pytest
from scipy import stats
from scipy.integrate import quad

import oracle

F_EPS = 1e-4


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_textbook_matches_bruteforce(n, seed):
    rng = np.random.default_rng(seed * 100 + n)
    nin = rng.uniform(0.05, 0.95, size=n)
    s = np.sort(rng.uniform(0.5, 2.0, size=n))
    d = float(rng.uniform(0.5, 2.0))
    sig = [0.1, 0.0, 0.0]
    psi = np.array([oracle.depth_psi(si, d, sig) for si in s])
    psi_inf = float(rng.uniform(0.0, 2.0))
    p, T = oracle.ray_messages(nin, np.ones(n), s, d, sig, mode=1, psi_inf=psi_inf)
    ref = oracle.ray_messages_bruteforce(nin, psi, psi_inf)
    np.testing.assert_allclose(p, ref, rtol=1e-12, atol=1e-14)
    # T is the factor's partition sum under the incoming messages
    Z = 0.0
    for o in range(1 << n):
        first = next((k for k in range(n) if o >> k & 1), None)
        w = psi_inf if first is None else psi[first]
        for j in range(n):
            w *= (1 - nin[j]) if o >> j & 1 else nin[j]
        Z += w
    assert T == pytest.approx(Z, rel=1e-12)


def test_bruteforce_detects_a_planted_bug():
    """A plausible mistake (visibility including the voxel itself, the
    reference's Q1 behaviour) is caught by the brute-force comparison."""
    rng = np.random.default_rng(7)
    n = 6
    nin = rng.uniform(0.1, 0.9, size=n)
    s = np.linspace(0.8, 1.2, n)
    psi = np.array([oracle.depth_psi(si, 1.0, [0.1, 0, 0]) for si in s])
    ref = oracle.ray_messages_bruteforce(nin, psi, 1.0)
    p_ref_mode, _ = oracle.ray_messages(nin, np.ones(n), s, 1.0, [0.1, 0, 0], mode=0, ran_out=True)
    assert np.abs(p_ref_mode - ref).max() > 1e-3


def test_psi_is_scaled_normal_pdf():
    sig, bia = [0.02, 0.001, 0.003], [0.01, -0.002, 0.0]
    for s in [0.5, 1.3, 2.7]:
        for d in [s - 0.1, s, s + 0.05]:
            mean = s + bia[0] + bia[1] * s + bia[2] * s * s
            sd = 3 * (sig[0] + sig[1] * s + sig[2] * s * s)
            want = 0.398942 * math.sqrt(2 * math.pi) * stats.norm.pdf(d, loc=mean, scale=sd)
            assert oracle.depth_psi(s, d, sig, bia) == pytest.approx(want, rel=1e-12)
    total, _ = quad(lambda d: oracle.depth_psi(1.0, d, [0.015, 0, 0]), -1, 3, points=[1.0], limit=200)
    assert total == pytest.approx(0.398942 * math.sqrt(2 * math.pi), rel=1e-9)


def test_linear_polynomial_degree_two():
    assert oracle.depth_psi(1.0, 1.0, [0.01, 0.02, 99.0], [0, 0, 0], lookup_n=2) == pytest.approx(
        0.398942 / (3 * 0.03), rel=1e-12)


def _psi(s, d, sg):
    u = abs(d - s) / (3 * sg)
    return 0.398942 * math.exp(-0.5 * u * u) / (3 * sg)


def test_reference_mode_one_voxel_by_hand():
    n0, l, s, d, sg = 0.7, 0.8, 1.0, 1.02, 0.01
    V = n0 ** l                       # visibility includes the voxel itself
    psi = _psi(s, d, sg)
    tau = V * (1 - n0) * psi
    # walk truncated (no infinity term): R = 0
    p, T = oracle.ray_messages([n0], [l], [s], d, [sg, 0, 0], ran_out=False)
    assert T == pytest.approx(tau, rel=1e-12)
    mu1, mu0 = V * psi, 0.0
    assert p[0] == pytest.approx(min(mu1 / (mu1 + mu0 + 1e-7), 1 - F_EPS), rel=1e-12)
    # ran out with d beyond the voxel: T gains V, R_1 = V
    p2, T2 = oracle.ray_messages([n0], [l], [s], d, [sg, 0, 0], ran_out=True)
    assert T2 == pytest.approx(tau + V, rel=1e-12)
    want = (V * psi) / (V * psi + V / n0 + 1e-7)
    assert p2[0] == pytest.approx(want, rel=1e-12)
    # ran out but d in front of the last voxel: no infinity term
    p3, T3 = oracle.ray_messages([n0], [l], [s], 0.98, [sg, 0, 0], ran_out=True)
    assert T3 == pytest.approx(V * (1 - n0) * _psi(s, 0.98, sg), rel=1e-12)


def test_reference_mode_two_voxels_by_hand():
    n = [0.9, 0.4]
    l = [1.0, 0.5]
    s = [0.99, 1.0]
    d, sg = 1.0, 0.01
    V1 = n[0] ** l[0]
    V2 = V1 * n[1] ** l[1]
    psi = [_psi(s[0], d, sg), _psi(s[1], d, sg)]
    t1, t2 = V1 * (1 - n[0]) * psi[0], V2 * (1 - n[1]) * psi[1]
    T = t1 + t2 + V2                  # ran out, d > s_n
    mu1 = [V1 * psi[0], t1 + V2 * psi[1]]
    mu0 = [(t2 + V2) / n[0], t1 + V2 / n[1]]
    want = [mu1[i] / (mu1[i] + mu0[i] + 1e-7) for i in range(2)]
    p, Tg = oracle.ray_messages(n, l, s, d + 1e-9, [sg, 0, 0], ran_out=True)
    assert Tg == pytest.approx(T, rel=1e-6)
    np.testing.assert_allclose(p, want, rtol=1e-6)


def test_clamp_and_normaliser():
    # all-zero likelihood: mu1 = mu0 = 0 -> 0 / 1e-7 -> clamped to F_EPS
    p, T = oracle.ray_messages([0.5, 0.5], [1, 1], [1.0, 1.01], 50.0, [0.001, 0, 0], ran_out=False)
    assert T == 0.0
    np.testing.assert_array_equal(p, F_EPS)


def test_free_space_then_surface():
    """A ray through prior-0.1 voxels ending on a surface: voxels well in
    front of the surface get P(occ) below the prior, the surface voxel gets
    the largest message and it exceeds 0.5."""
    res, n = 0.04, 60
    s = np.arange(1, n + 1) * res
    d = 2.0
    p, _ = oracle.ray_messages(np.full(n, 0.9), np.ones(n), s, d, [1.5 * res, 0, 0], ran_out=False)
    assert np.all((p >= F_EPS) & (p <= 1 - F_EPS))
    front = s < d - 0.6             # more than 3.3 likelihood widths (3 sigma) in front
    assert np.all(p[front] < 0.05)
    k = int(np.argmax(p))
    assert abs(s[k] - d) <= 0.2 and p[k] > 0.5


def test_gap_records_do_not_attenuate():
    """Gap records (live = 0) advance depth only: inserting one changes no
    message of the live voxels."""
    nin = np.array([0.8, 0.6, 0.3])
    s = np.array([0.95, 1.0, 1.05])
    p, T = oracle.ray_messages(nin, [1, 1, 1], s, 1.0, [0.02, 0, 0], ran_out=False)
    nin2 = np.array([0.8, 1.0, 0.6, 0.3])
    s2 = np.array([0.95, 0.97, 1.0, 1.05])
    p2, T2 = oracle.ray_messages(nin2, [1, 5, 1, 1], s2, 1.0, [0.02, 0, 0], live=[1, 0, 1, 1], ran_out=False)
    assert T2 == pytest.approx(T, rel=1e-14)
    np.testing.assert_allclose(p2[[0, 2, 3]], p, rtol=1e-14)


def test_set_pose_resets_voxels_no_other_camera_reaches():
    """set_pose (GVDBInference.cpp:66-82): after the moved camera's rays are
    removed, voxels whose summed chord length over the other cameras is below
    0.05 return to the prior and read 0 (reset_empty_voxels,
    gvdbkernels.cu:498-517); voxels the other cameras still reach keep their
    beliefs."""
    from mrfmap_amd import synth
    from helpers import oracle_for

    p, frames, _ = synth.make_workload("small", frames=2)
    o = oracle_for(p, frames)
    o.perform_inference(2)
    B0 = o.beliefs().ravel().copy()
    prior = np.float64(np.log(np.float32(p.prior)) - np.log(np.float32(1 - p.prior)))
    # voxels the second camera reaches, from its traversal
    _, rec = o.traversal(1)
    vox = rec[:, 0][rec[:, 0] != 0xFFFFFFFF]
    length = rec[:, 1][rec[:, 0] != 0xFFFFFFFF].view(np.float32).astype(np.float64)
    cum = np.bincount(vox, weights=length, minlength=B0.size)
    moved = frames[0][0].copy()
    moved[0, 3] += 0.1
    o.set_pose(0, moved)
    # the reset is immediate; reading re-traverses the moved camera, whose new
    # bricks are appended after the existing ones
    B1 = o.beliefs().ravel()[:B0.size]
    reset = cum < 0.05
    np.testing.assert_array_equal(B1[reset], np.full(reset.sum(), prior))
    np.testing.assert_array_equal(B1[~reset], B0[~reset])
    assert reset.sum() > 0 and (~reset).sum() > 0
    a = o.alphas().ravel()[:B0.size]
    assert (a[reset] == 0).all()


def test_set_pose_single_camera_clears_the_map():
    from mrfmap_amd import synth
    from helpers import oracle_for

    p, frames, _ = synth.make_workload("small", frames=1)
    o = oracle_for(p, frames)
    o.perform_inference(2)
    assert (o.alphas() != 0).any()
    o.set_pose(0, frames[0][0])
    assert (o.alphas() == 0).all()
