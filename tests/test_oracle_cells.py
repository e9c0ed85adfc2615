"""Pins the oracle's per-cell sensor model (the use_polys tables of
GVDBParams.h:93-104, 134-162, looked up per pixel at gvdbkernels.cu:46-55).

A table cell covers 20 x 20 pixels of the vertically flipped image and holds
lookup_n coefficients: pixel (x, y) of a rows x cols image reads cell
((rows - 1 - y) / 20) * (n_u * lookup_n) + (x / 20) * lookup_n, n_u =
ceil(cols / 20).  The closed-form check: when only the pixels of one cell
carry depth, changing the coefficients of exactly that cell changes the
traversal and the map, and changing any other cell (the unflipped row, the
transposed cell, a neighbour) changes nothing, bit for bit.
"""
import numpy as np
import pytest

import oracle
from mrfmap_amd import synth

ROWS, COLS = 120, 160
NU, NV = 8, 6  # cells per row and per column


def _scene(depth_cell_xy):
    """One 160x120 keyframe of the small workload whose valid depth is
    limited to the pixels of one 20x20 block (block column bx, block row by
    of the unflipped image)."""
    p, frames, _ = synth.make_workload("small", frames=1)
    pose, K, depth = frames[0]
    bx, by = depth_cell_xy
    d = np.zeros_like(depth)
    d[by * 20:(by + 1) * 20, bx * 20:(bx + 1) * 20] = depth[by * 20:(by + 1) * 20, bx * 20:(bx + 1) * 20]
    assert (d > 0).sum() > 300
    return p, pose, K, d


def _tables(p, ln, odd=None):
    sig = np.tile(np.asarray([1.5 * p.res, 0.004, 0.001][:ln], np.float32), NU * NV)
    bia = np.tile(np.asarray([0.003, -0.001, 0.0005][:ln], np.float32), NU * NV)
    if odd is not None:
        row, col = odd
        o = (row * NU + col) * ln
        sig[o:o + ln] *= 2.0
        bia[o] += 0.02
    return sig, bia


def _run(p, pose, K, d, ln, sig, bia, iters=2):
    o = oracle.Oracle(p.K, ROWS, COLS, p.res, p.prior, p.dims, lookup_n=ln, activation=1, sigma_lookup=sig,
                      bias_lookup=bia)
    o.add_camera(pose, d, K)
    o.perform_inference(iters)
    rays, recs = o.traversal(0)
    return rays, recs, o.alphas()


@pytest.mark.parametrize("ln", [2, 3])
def test_only_the_pixels_own_cell_matters(ln):
    # pixels x in [40, 60), y in [0, 20): flipped rows 100..119 -> cell row 5, column 2
    p, pose, K, d = _scene((2, 0))
    base = _run(p, pose, K, d, ln, *_tables(p, ln))
    rays = base[0]
    np.testing.assert_array_equal(rays[:, 5] & 0x7FFFFFFF, (5 * NU + 2) * ln)
    own = _run(p, pose, K, d, ln, *_tables(p, ln, odd=(5, 2)))
    assert own[1].shape != base[1].shape or (own[1] != base[1]).any()  # wider sigma: later truncation
    assert own[2].shape != base[2].shape or np.abs(own[2] - base[2]).max() > 1e-3
    for other in [(0, 2), (2, 5), (5, 3), (4, 2)]:  # unflipped row, transposed, neighbours
        r = _run(p, pose, K, d, ln, *_tables(p, ln, odd=other))
        np.testing.assert_array_equal(r[0], base[0])
        np.testing.assert_array_equal(r[1], base[1])
        np.testing.assert_array_equal(r[2], base[2])


@pytest.mark.parametrize("ln", [2, 3])
def test_constant_tables_equal_constant_polynomials(ln):
    """Tables holding the same polynomial in every cell give the result of
    the constant polynomials, bit for bit."""
    p, frames, _ = synth.make_workload("small", frames=2)
    sp, bp = [1.5 * p.res, 0.004, 0.001][:ln], [0.003, -0.001, 0.0005][:ln]
    sig, bia = _tables(p, ln)
    a = oracle.Oracle(p.K, ROWS, COLS, p.res, p.prior, p.dims, sp, bp, lookup_n=ln, activation=1)
    b = oracle.Oracle(p.K, ROWS, COLS, p.res, p.prior, p.dims, lookup_n=ln, activation=1, sigma_lookup=sig,
                      bias_lookup=bia)
    for pose, K, depth in frames:
        a.add_camera(pose, depth, K)
        b.add_camera(pose, depth, K)
    a.perform_inference(2)
    b.perform_inference(2)
    for c in range(2):
        ta, tb = a.traversal(c), b.traversal(c)
        np.testing.assert_array_equal(ta[0], tb[0])
        np.testing.assert_array_equal(ta[1], tb[1])
    np.testing.assert_array_equal(a.alphas(), b.alphas())


def test_tables_require_the_params_image_size():
    p, frames, _ = synth.make_workload("small", frames=1)
    sig, bia = _tables(p, 3)
    o = oracle.Oracle(p.K, ROWS, COLS, p.res, p.prior, p.dims, lookup_n=3, activation=1, sigma_lookup=sig,
                      bias_lookup=bia)
    with pytest.raises(ValueError):
        o.add_camera(frames[0][0], np.ones((ROWS, COLS + 8), np.float32), frames[0][1])
