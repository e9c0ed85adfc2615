"""Pins the FP64 warm-start MLP of the oracle (oracle/mlp.py).

* the backward pass of the plain FP64 network against central finite
  differences of its own loss, on parameters of every block;
* the forward pass against closed forms: with W1 = 0 the output does not
  depend on x and equals b3 + w3 . gelu(b2 + W2 gelu(b1)) evaluated by hand
  for constant biases; GELU at 0 and far from 0, and its derivative
  against finite differences;
* BF16 rounding (nearest, ties to even) on hand-picked bit patterns.
"""
import math

import numpy as np
import pytest

from oracle import mlp


def _problem(n=24, seed=3):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed).astype(np.float64)
    p[mlp.B1] = rng.normal(0, 0.3, mlp.HID)
    p[mlp.B2] = rng.normal(0, 0.3, mlp.HID)
    p[mlp.B3] = 0.2
    X = rng.normal(0, 1, (n, mlp.IN))
    t = rng.normal(0, 2, n)
    return p, X, t


def test_gradient_matches_finite_differences():
    p, X, t = _problem()
    _, g = mlp.loss_and_grad(p, X, t, bf16_ops=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(np.arange(s.start, s.stop), size=min(12, s.stop - s.start), replace=False)
                          for s in (mlp.W1, mlp.B1, mlp.W2, mlp.B2, mlp.W3, mlp.B3)])
    h = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (mlp.loss_and_grad(pp, X, t, bf16_ops=False)[0] - mlp.loss_and_grad(pm, X, t, bf16_ops=False)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def test_forward_closed_form_without_input_weights():
    p = np.zeros(mlp.NPARAMS)
    p[mlp.B1] = 0.5
    W2 = np.zeros((mlp.HID, mlp.HID))
    W2[:, 0] = 1.0  # every layer-2 unit sees hidden unit 0 only
    p[mlp.W2] = W2.ravel()
    p[mlp.B2] = -0.25
    p[mlp.W3] = 0.125
    p[mlp.B3] = 1.5
    X = np.random.default_rng(0).normal(0, 3, (5, mlp.IN))
    y, _ = mlp.forward(p, X, bf16_ops=False)
    g1 = 0.5 * 0.5 * (1 + math.erf(0.5 / math.sqrt(2)))
    a2 = g1 - 0.25
    g2 = a2 * 0.5 * (1 + math.erf(a2 / math.sqrt(2)))
    np.testing.assert_allclose(y, 1.5 + 64 * 0.125 * g2, rtol=1e-14)


def test_gelu_values_and_derivative():
    assert mlp.gelu(np.array(0.0)) == 0.0
    assert mlp.gelu_d(np.array(0.0)) == 0.5
    np.testing.assert_allclose(mlp.gelu(np.array([10.0, -10.0])), [10.0, 0.0], atol=1e-20)
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_d(z), fd, atol=1e-8)


@pytest.mark.parametrize("bits,want", [
    (0x3F800000, 0x3F80),  # 1.0 exact
    (0x3F808000, 0x3F80),  # halfway, even below -> down
    (0x3F818000, 0x3F82),  # halfway, odd below -> up
    (0x3F808001, 0x3F81),  # above halfway -> up
    (0xBF807FFF, 0xBF80),  # below halfway, negative -> toward zero
    (0x7F7FFFFF, 0x7F80),  # largest float rounds to infinity
])
def test_bf16_rounding(bits, want):
    x = np.array([bits], dtype=np.uint32).view(np.float32)
    r = mlp.bf16(x)
    got = (np.float32(r[0]).view(np.uint32) >> 16) & 0xFFFF
    assert got == want
    assert (np.float32(r[0]).view(np.uint32) & 0xFFFF) == 0


def test_bf16_gradient_differs_only_by_rounding():
    """The BF16 variant rounds inputs of the products only: with parameters
    and inputs already on the BF16 grid its forward pass matches the plain
    network up to the rounding of h1."""
    p, X, t = _problem(n=8)
    p = mlp.bf16(p)
    X = mlp.bf16(X)
    y64, _ = mlp.forward(p, X, bf16_ops=False)
    y16, _ = mlp.forward(p, X, bf16_ops=True)
    assert np.max(np.abs(y64 - y16)) < 0.05 * (1 + np.max(np.abs(y64)))
    assert not np.array_equal(y64, y16)
