"""The OctoMap-equivalent CPU comparator (oracle/octomap_eq.c): hit/miss
log-odds, clamping, free-space ray keys and the BBX limit of
GVDBOctomapWrapper::init (src/GVDBOctomapWrapper.cpp:3-17)."""
import math

import numpy as np
import pytest

import oracle

L_HIT, L_MISS = math.log(0.7 / 0.3), math.log(0.4 / 0.6)


def test_single_ray_hit_and_misses():
    m = oracle.OctomapEq(0.1, (4.0, 4.0, 3.0))
    n = m.insert_point_cloud(np.array([[1.05, 0.05, 1.05]]), [0.05, 0.05, 1.05])
    assert n == 11  # 10 free keys + 1 occupied
    for i in range(10):
        assert m.logodds_at([0.05 + 0.1 * i, 0.05, 1.05]) == pytest.approx(L_MISS, rel=1e-6)
    assert m.logodds_at([1.05, 0.05, 1.05]) == pytest.approx(L_HIT, rel=1e-6)
    assert math.isnan(m.logodds_at([1.15, 0.05, 1.05]))


def test_occupied_wins_within_a_scan_and_clamps():
    m = oracle.OctomapEq(0.1, (4.0, 4.0, 3.0))
    pts = np.array([[0.55, 0.05, 1.05], [1.05, 0.05, 1.05]])  # second ray passes the first endpoint
    for _ in range(40):
        m.insert_point_cloud(pts, [0.05, 0.05, 1.05])
    assert m.logodds_at([0.55, 0.05, 1.05]) == pytest.approx(math.log(0.9999 / 0.0001), rel=1e-6)
    assert m.logodds_at([0.25, 0.05, 1.05]) == pytest.approx(math.log(0.0001 / 0.9999), rel=1e-6)


def test_bbx_limits_updates():
    m = oracle.OctomapEq(0.1, (4.0, 4.0, 3.0))
    # origin above the BBX (z > 3): the first key is outside, so no free space
    m.insert_point_cloud(np.array([[0.05, 0.05, 2.05]]), [0.05, 0.05, 3.55])
    assert math.isnan(m.logodds_at([0.05, 0.05, 2.55]))
    assert m.logodds_at([0.05, 0.05, 2.05]) == pytest.approx(L_HIT, rel=1e-6)
