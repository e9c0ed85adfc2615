"""mrf_mlp_grad_noalloc, the entry point that allocates nothing, against
mrf_mlp_grad: device buffers in and out, per-wave partial gradients in the
library's fixed workspace.  Both run the same kernels with the same wave
count, so gradients and loss must agree bit for bit; two calls in a row,
the second smaller (the workspace keeps the first call's contents), and a
size past the workspace is refused before anything is launched.

Device buffers are allocated here with the HIP runtime that libmrfmap_hip.so
itself links (looked up by soname), not through torch.
"""
import ctypes

import numpy as np
import pytest

from mrfmap_amd import _lib, synth

import oracle.mlp as omlp
from helpers import gpu_for

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.load()  # binds libamdhip64.so.7 first, so the lookup below returns the same runtime
    hip = ctypes.CDLL("libamdhip64.so.7")
    P = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
    hip.hipFree.argtypes = [P]
    hip.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(P)]
    hip.hipStreamDestroy.argtypes = [P]
    hip.hipStreamSynchronize.argtypes = [P]
    return hip


class _Dev:
    """A device copy of a numpy array (or zeros of a shape and dtype)."""

    def __init__(self, hip, arr):
        self.hip = hip
        self.arr = np.ascontiguousarray(arr)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(self.arr.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.arr.ctypes.data_as(ctypes.c_void_p), self.arr.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.arr)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _params(seed):
    return (omlp.init_params(seed) * 1.5).astype(np.float32)


def _noalloc(g, hip, params, X, t, stream, with_loss=True):
    bufs = [_Dev(hip, params), _Dev(hip, X), _Dev(hip, t), _Dev(hip, np.zeros(_lib.MLP_PARAMS, np.float32)),
            _Dev(hip, np.full(1, 123.0, np.float64))]
    try:
        dp, dx, dt, dg, dl = bufs
        assert hip.hipDeviceSynchronize() == 0  # the copies are done before any stream reads them
        rc = g._L.mrf_mlp_grad_noalloc(g._h, dp.ptr, dx.ptr, dt.ptr, len(X), dg.ptr, dl.ptr if with_loss else None,
                                       stream)
        _lib.check(rc, g._h)
        assert hip.hipDeviceSynchronize() == 0
        return float(dl.get()[0]), dg.get()
    finally:
        for b in bufs:
            b.free()


def test_noalloc_gradient_matches_mlp_grad():
    p, frames, _ = synth.make_workload("small", frames=1)
    g = gpu_for(p, frames)
    hip = _hip()
    params = _params(7)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # the largest size the workspace holds, then a smaller ragged one on the
        # context's own stream: the second call must not see the first's partials
        for i, (n, s) in enumerate(((_lib.MLP_NOALLOC_MAX_SAMPLES, stream), (1000 + 13, None))):
            rng = np.random.default_rng(30 + i)
            X = rng.normal(0, 1, (n, 32)).astype(np.float32)
            t = rng.normal(0, 2, n).astype(np.float32)
            print(f"call {i}: n={n} stream={'own' if s else 'context'}", flush=True)
            loss_ref, grad_ref = g.mlp_grad(params, X, t)
            loss, grad = _noalloc(g, hip, params, X, t, s)
            np.testing.assert_array_equal(grad.view(np.uint32), grad_ref.view(np.uint32))
            assert loss == loss_ref, (loss, loss_ref)
            assert np.abs(grad).max() > 0
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_without_loss_and_limits():
    p, frames, _ = synth.make_workload("small", frames=1)
    g = gpu_for(p, frames)
    hip = _hip()
    params = _params(8)
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (77, 32)).astype(np.float32)
    t = rng.normal(0, 2, 77).astype(np.float32)
    _, grad_ref = g.mlp_grad(params, X, t)
    loss, grad = _noalloc(g, hip, params, X, t, None, with_loss=False)
    assert loss == 123.0  # the loss buffer was not passed, so it is untouched
    np.testing.assert_array_equal(grad.view(np.uint32), grad_ref.view(np.uint32))
    # past the workspace: refused with MRF_E_LIMIT before any launch (no device buffers needed)
    dummy = ctypes.c_void_p(8)
    rc = g._L.mrf_mlp_grad_noalloc(g._h, dummy, dummy, dummy, _lib.MLP_NOALLOC_MAX_SAMPLES + 1, dummy, None, None)
    assert rc == -7
    assert b"workspace" in g._L.mrf_last_error(g._h)
    rc = g._L.mrf_mlp_grad_noalloc(g._h, None, dummy, dummy, 10, dummy, None, None)
    assert rc == -1
