"""CPU-side checks of the boundary: exported symbols, loud failure without a
GPU, the gvdb_params mirror and the recorded-frame fixtures."""
import ctypes
import os
import re

import numpy as np
import pytest

from mrfmap_amd import GVDBInference, _lib, gvdb_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _header_symbols(name):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mrf_[a-z0-9_]+)\s*\(", txt)))


def test_hip_library_exports_every_declared_symbol():
    syms = _header_symbols("mrfmap_hip.h")
    assert len(syms) >= 14
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms


def test_synth_library_exports_every_declared_symbol():
    from mrfmap_amd import synth
    syms = _header_symbols("mrfmap_synth.h")
    L = ctypes.CDLL(synth.LIB_PATH)
    assert not [s for s in syms if not hasattr(L, s)]


def _has_gpu():
    try:
        L = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int()
        return L.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a HIP device is present")
def test_no_device_fails_loudly():
    """No CPU fallback: without a HIP device creating a context raises."""
    p = gvdb_params()
    with pytest.raises(_lib.MrfError) as e:
        GVDBInference(params=p)
    assert e.value.code == -2


def test_params_defaults_follow_reference():
    """gvdb_params_t() defaults (GVDBParams.h:32-55)."""
    p = gvdb_params()
    assert p.res == 0.01 and p.prior == 0.01 and p.max_iters == 1
    assert p.occ_thresh == 0.1 and p.acc_thresh == 0.5
    assert p.lookup_n == 3 and p.const_sigmasq_poly == [0.015, 0.0, 0.0]
    c = _lib.MrfParams()
    L = _lib.load()
    L.mrf_params_default(ctypes.byref(c))
    assert c.res == pytest.approx(0.01) and c.prior == pytest.approx(0.01) and c.max_iters == 1
    assert c.const_sigmasq_poly[0] == pytest.approx(0.015)
    assert c.K[0] == pytest.approx(p.K[0, 0]) and c.K[2] == pytest.approx(0.5)


def test_params_load_from_file():
    p = gvdb_params().load_from_file(os.path.join(GOLD, "coffin_world_640.yaml"))
    assert p.rows == 480 and p.cols == 640
    assert p.res == pytest.approx(0.02) and p.prior == pytest.approx(0.1)
    np.testing.assert_allclose(p.dims, [4, 4, 3])
    assert p.max_iters == 3 and p.lookup_n == 3
    assert p.const_sigmasq_poly == [0.05, 0.0, 0.0]
    assert p.K[0, 0] == pytest.approx(381.36115025) and p.K[1, 2] == pytest.approx(239.5)
    # one polynomial per 20x20 cell: 32 x 24 cells x 3 coefficients
    assert p.depth_sigma_lookup.shape == (32 * 24 * 3,)
    assert p.gvdb_map_config == [3, 3, 3, 3, 3]


def test_params_missing_node_raises(tmp_path):
    f = tmp_path / "bad.yaml"
    f.write_text("other: {}\n")
    with pytest.raises(ValueError):
        gvdb_params().load_from_file(str(f))


def test_recorded_frame_fixture():
    """data/depth_k.bin of the reference is noise-free planar depth: the bottom
    row sees the floor at constant z (1.9159 m) and the top rows are NaN."""
    z = np.load(os.path.join(GOLD, "coffin_frames.npz"))
    depth, pose = z["depth"], z["body_pose"]
    assert depth.shape == (2, 480, 640)
    for d in depth:
        assert np.isnan(d).mean() == pytest.approx(0.225, abs=1e-3)
        np.testing.assert_allclose(d[479], 1.9159354, rtol=1e-6)
    for T in pose:
        assert np.hypot(T[0, 3], T[1, 3]) == pytest.approx(3.0, abs=1e-5)
        assert T[2, 3] == pytest.approx(2.0)
        R = T[:3, :3]
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)


CPP_CALLER = r"""
#include <cstdio>
#include "mrfmap_hip.hpp"
int main() {
  mrfmap_hip::Params p;
  p.res = 0.04f;
  std::printf("defaults res=%.2f prior=%.2f max_iters=%d lookup_n=%d\n", p.res, p.prior, p.max_iters, p.lookup_n);
  try {
    mrfmap_hip::GVDBInference inf(p);
    std::printf("created cameras=%u bricks=%lld\n", inf.num_cameras(), (long long)inf.num_bricks());
  } catch (const mrfmap_hip::Error& e) {
    std::printf("error %d\n", e.code());
  }
  return 0;
}
"""


def test_cpp_wrapper_compiles_links_and_reports(tmp_path):
    """include/mrfmap_hip.hpp (the GVDBInference-named C++ surface) builds
    with the host compiler against libmrfmap_hip.so; without a device the
    constructor throws MRF_E_HIP, with one it creates an empty context."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "caller.cpp"
    src.write_text(CPP_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lmrfmap_hip", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    assert "defaults res=0.04 prior=" in out
    if _has_gpu():
        assert "created cameras=0 bricks=0" in out
    else:
        assert "error -2" in out
