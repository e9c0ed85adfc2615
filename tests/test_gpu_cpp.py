"""The C++ surface (include/mrfmap_hip.hpp, GVDBInference's method names) on
a device: tests/cpp/run_workload, built with the package, runs the small
workload through Params / GVDBInference / add_camera_with_depth /
perform_inference / get_alphas / get_occupied_bricks.  Its marginals equal
the Python path's bit for bit (the same library with the same parameters)
and lie within 1e-3 of the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import gpu_for, oracle_for
from mrfmap_amd import synth

RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "run_workload")


def _write_workload(path, p, frames):
    rows, cols = p.rows, p.cols
    with open(path, "wb") as f:
        np.array([len(frames), rows, cols, p.max_iters, p.activation, p.lookup_n, p.brick_local_vis],
                 dtype=np.int32).tofile(f)
        np.array([p.res, p.prior, *p.dims, *p.const_sigmasq_poly, *p.const_bias_poly], dtype=np.float32).tofile(f)
        for pose, K, depth in frames:
            np.asarray(pose, np.float32).tofile(f)
            np.asarray(K, np.float32).tofile(f)
            np.ascontiguousarray(depth, np.float32).tofile(f)


def _read_result(path):
    with open(path, "rb") as f:
        nb = int(np.fromfile(f, np.int64, 1)[0])
        a = np.fromfile(f, np.float32, nb * 512).reshape(nb, 512)
        centres = np.fromfile(f, np.float32, nb * 3).reshape(nb, 3)
        iters = int(np.fromfile(f, np.int32, 1)[0])
    return a, centres, iters


@pytest.mark.gpu
def test_cpp_surface_runs_inference_on_device(tmp_path):
    if not os.path.exists(RUNNER):
        pytest.fail("tests/cpp/run_workload is not built (python -m mrfmap_amd.build)")
    p, frames, _ = synth.make_workload("small")
    p.max_iters = 3
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_workload(src, p, frames)
    print(f"C++ runner: {len(frames)} keyframes {p.rows}x{p.cols}, {p.max_iters} iterations", flush=True)
    r = subprocess.run([RUNNER, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    a, centres, iters = _read_result(dst)
    assert iters == 3
    g = gpu_for(p, frames)
    g.perform_inference(max_iters=3)
    np.testing.assert_array_equal(a.view(np.uint32), g.get_alphas().view(np.uint32))
    np.testing.assert_array_equal(centres, g.get_occupied_bricks())
    o = oracle_for(p, frames)
    o.perform_inference(3)
    err = np.abs(a.astype(np.float64) - o.alphas()).max()
    print(f"C++ surface: {a.shape[0]} bricks, max-abs marginal error {err:.3e}")
    assert err <= 1e-3
