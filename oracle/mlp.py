"""FP64 restatement of the warm-start MLP — TEST INFRASTRUCTURE ONLY.

The network of the learned warm start (mrfmap_amd/csrc/warm.hip; it has no
counterpart in the reference): 32 inputs -> 64 -> 64 -> 1, GELU
(z Phi(z), erf form), parameters as one float vector of 6337 values
[W1 64x32 | b1 64 | W2 64x64 | b2 64 | w3 64 | b3 1], matrices row-major.
Loss 1/(2n) sum (y - t)^2.

With ``bf16=True`` the values the device feeds its BF16 matrix cores are
rounded to BF16 (nearest even) exactly where the device rounds them: x, W1,
W2 and h1 in the forward pass, dz2 and dz1 in the backward pass; everything
else is FP64 here (FP32 on the device).  With ``bf16=False`` it is the plain
FP64 network, whose gradient the tests pin against central finite
differences.
"""
import numpy as np

IN, HID = 32, 64
W1 = slice(0, HID * IN)
B1 = slice(W1.stop, W1.stop + HID)
W2 = slice(B1.stop, B1.stop + HID * HID)
B2 = slice(W2.stop, W2.stop + HID)
W3 = slice(B2.stop, B2.stop + HID)
B3 = slice(W3.stop, W3.stop + 1)
NPARAMS = B3.stop  # 6337

_SQ2 = np.sqrt(2.0)


def _erf(x):
    from scipy.special import erf

    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / _SQ2))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / _SQ2)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def bf16(a):
    """Round float values to the nearest BF16 (ties to even) via their FP32 bits."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1].reshape(HID, IN), p[B1], p[W2].reshape(HID, HID), p[B2], p[W3], p[B3][0])


def init_params(seed=0):
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    p[W1] = rng.uniform(-1, 1, HID * IN) * np.sqrt(6.0 / IN)
    p[W2] = rng.uniform(-1, 1, HID * HID) * np.sqrt(6.0 / HID)
    p[W3] = rng.uniform(-1, 1, HID) * np.sqrt(6.0 / HID)
    return p


def forward(p, X, bf16_ops=True):
    """y (n,) and the intermediates of every sample."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    xb, w1b, w2b = r(X), r(w1), r(w2)
    a1 = xb @ w1b.T + b1
    h1 = r(gelu(a1))
    a2 = h1 @ w2b.T + b2
    h2 = gelu(a2)
    y = h2 @ w3 + b3
    return y, dict(xb=xb, w2b=w2b, a1=a1, h1=h1, a2=a2, h2=h2)


def loss_and_grad(p, X, t, bf16_ops=True):
    """(loss, gradient) of 1/(2n) sum (y - t)^2 in the parameter layout."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    y, c = forward(p, X, bf16_ops)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    dy = (y - t) / n
    loss = 0.5 * np.sum((y - t) ** 2) / n
    g = np.zeros(NPARAMS, np.float64)
    g[W3] = c["h2"].T @ dy
    g[B3] = dy.sum()
    dz2 = (dy[:, None] * w3[None, :]) * gelu_d(c["a2"])
    g[B2] = dz2.sum(0)
    dz2b = r(dz2)
    g[W2] = (dz2b.T @ c["h1"]).ravel()
    dh1 = dz2b @ c["w2b"]
    dz1 = dh1 * gelu_d(c["a1"])
    g[B1] = dz1.sum(0)
    g[W1] = (r(dz1).T @ c["xb"]).ravel()
    return loss, g
