"""Pins the oracle's traversal (orc_walk_ray) independently of it.

* exact brute force: an independent Python walk over exact Fractions
  (every voxel-boundary crossing inside the grid, merged, voxel of each
  interval from its midpoint) must give the same voxels and, through the
  contract's IEEE double formula, the same float chord lengths;
* closed forms: axis-aligned rays have l = 1 per interior voxel, a ray
  through voxel corners along a 45 degree diagonal has l = sqrt(2);
* invariants: sum of l equals the ray length inside the grid; exit depths
  increase; consecutive voxels are face/edge/corner neighbours;
* truncation: the walk stops at the first brick exit with
  z + bias(z) >= d + 3 sigma(z) (src/cuda/gvdbkernels.cu:567-570).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

FIX = 1 << 16


def fixed(v):
    return int(round(v * FIX))  # Python round = round-half-even like rint


def brute_walk(O, D, N):
    """Independent exact walk: returns [(voxel, z0, z1)] with Fractions."""
    Oi = [fixed(v) for v in O]
    Di = [fixed(v) for v in D]
    zin, zout = Fraction(0), None
    for a in range(3):
        top = N[a] * FIX
        if Di[a] == 0:
            if not (0 <= Oi[a] < top):
                return [], Oi, Di
            continue
        lo = Fraction(0 - Oi[a], Di[a]) if Di[a] > 0 else Fraction(top - Oi[a], Di[a])
        hi = Fraction(top - Oi[a], Di[a]) if Di[a] > 0 else Fraction(0 - Oi[a], Di[a])
        zin = max(zin, lo)
        zout = hi if zout is None else min(zout, hi)
    if zout is None or not zin < zout:
        return [], Oi, Di
    cuts = set()
    for a in range(3):
        if Di[a] == 0:
            continue
        for k in range(1, N[a]):
            z = Fraction(k * FIX - Oi[a], Di[a])
            if zin < z < zout:
                cuts.add(z)
    pts = [zin] + sorted(cuts) + [zout]
    out = []
    for z0, z1 in zip(pts[:-1], pts[1:]):
        mid = (z0 + z1) / 2
        vox = tuple(math.floor((Oi[a] + mid * Di[a]) / FIX) for a in range(3))
        out.append((vox, z0, z1))
    return out, Oi, Di


def contract_len(z0, z1, Di):
    """chord length as the traversal contract evaluates it (IEEE double)."""
    n2 = 0.0
    for a in range(3):
        dd = Di[a] / 65536.0
        n2 = n2 + dd * dd
    norm = math.sqrt(n2)
    zc = float(z0.numerator) / float(z0.denominator)
    zn = float(z1.numerator) / float(z1.denominator)
    return np.float32((zn - zc) * norm)


@pytest.mark.parametrize("seed", range(12))
def test_walk_matches_exact_bruteforce(seed):
    rng = np.random.default_rng(seed)
    N = [int(v) for v in rng.integers(5, 24, size=3)]
    O = rng.uniform(-4.0, np.array(N) + 4.0)
    D = rng.normal(size=3) * rng.uniform(2, 12)
    if seed % 4 == 0:
        D[rng.integers(3)] = 0.0  # axis-parallel component
        O = rng.uniform(0.5, np.array(N) - 0.5)
    vox, ln, zx, _ = oracle.walk_ray(O, D, N)
    ref, Oi, Di = brute_walk(O, D, N)
    assert len(vox) == len(ref)
    for (v, z0, z1), got_v, got_l, got_z in zip(ref, vox, ln, zx):
        assert tuple(got_v) == v
        assert np.float32(got_l) == contract_len(z0, z1, Di)
        assert got_z == float(z1.numerator) / float(z1.denominator)


def test_axis_aligned_unit_chords():
    N = [20, 10, 10]
    vox, ln, zx, ran_out = oracle.walk_ray([0.25, 3.5, 4.5], [10.0, 0.0, 0.0], N)
    assert [tuple(v) for v in vox] == [(i, 3, 4) for i in range(20)]
    assert ln[0] == pytest.approx(0.75)
    np.testing.assert_allclose(ln[1:], 1.0, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(ln[1:].astype(np.float32), np.float32(1.0))
    assert ran_out


def test_diagonal_through_corners_sqrt2():
    """Ties step both axes at once: voxels (i, i) with l = sqrt(2)."""
    N = [16, 16, 4]
    vox, ln, _, _ = oracle.walk_ray([0.0, 0.0, 1.5], [8.0, 8.0, 0.0], N)
    assert [tuple(v) for v in vox] == [(i, i, 1) for i in range(16)]
    np.testing.assert_allclose(ln, math.sqrt(2.0), rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_length_sum_and_adjacency(seed):
    rng = np.random.default_rng(100 + seed)
    N = [40, 40, 30]
    O = rng.uniform(0, np.array(N, dtype=float))
    D = rng.normal(size=3) * 50
    vox, ln, zx, _ = oracle.walk_ray(O, D, N)
    ref, Oi, Di = brute_walk(O, D, N)
    total = float(ref[-1][2] - ref[0][1]) * math.sqrt(sum((d / 65536.0) ** 2 for d in Di))
    assert ln.sum() == pytest.approx(total, rel=1e-12)
    assert np.all(np.diff(zx) > 0)
    steps = np.abs(np.diff(vox, axis=0))
    assert np.all(steps.max(axis=1) == 1)


def test_truncation_at_brick_exit():
    """Depth 1.0 m, sigma 0.05 -> stop after the first brick whose exit z
    satisfies z >= d + 0.15; voxel size 1 cm along +x at 100 voxels/m."""
    N = [256, 16, 16]
    d, sig = 1.0, 0.05
    vox, ln, zx, ran_out = oracle.walk_ray([0.5, 4.5, 4.5], [100.0, 0.0, 0.0], N, depth=d, trunc=True,
                                           sigma_poly=[sig, 0, 0])
    assert not ran_out
    last = vox[-1][0]
    assert (last + 1) % 8 == 0                      # ends on a brick boundary
    assert zx[-1] >= d + 3 * sig                    # condition holds at that exit
    prev_exit = zx[np.nonzero((vox[:, 0] + 1) % 8 == 0)[0]]
    assert np.all(prev_exit[:-1] < d + 3 * sig)     # and at no earlier brick exit


def test_truncation_ran_out():
    N = [64, 8, 8]
    vox, ln, zx, ran_out = oracle.walk_ray([0.5, 4.5, 4.5], [100.0, 0.0, 0.0], N, depth=5.0, trunc=True,
                                           sigma_poly=[0.05, 0, 0])
    assert ran_out and len(vox) == 64


def test_cum_lengths_are_the_traversal_chord_sums():
    """orc_get_cum_lengths (the CUM_LENS channel, gvdbkernels.cu:1276-1302)
    equals, voxel by voxel, the sum of the chord lengths of the camera's
    exported traversal records."""
    from helpers import oracle_for
    from mrfmap_amd import synth
    p, frames, _ = synth.make_workload("small", frames=2)
    o = oracle_for(p, frames)
    o.prepare()
    for c in range(len(frames)):
        _, recs = o.traversal(c)
        live = recs[:, 0] != 0xFFFFFFFF
        lens = recs[live, 1].view(np.float32).astype(np.float64)
        cum = o.cum_lengths(c).reshape(-1)
        expect = np.bincount(recs[live, 0].astype(np.int64), weights=lens, minlength=cum.size)
        assert cum.size == expect.size
        np.testing.assert_allclose(cum, expect, rtol=1e-12, atol=0)
        assert (cum > 0).sum() > 1000
