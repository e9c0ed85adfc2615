"""Build the native libraries in-tree.

libmrfmap_hip.so   HIP kernels + C ABI for gfx950 (hipcc)
libmrfmap_synth.so synthetic keyframe generator (host C, gcc)

Run ``python -m mrfmap_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("MRF_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["api.hip", "traverse.hip", "bp.hip", "scan.hip", "geometry.hip", "warm.hip"]
# traverse.hip: every double operation is one IEEE operation (bit-exact lists);
# bp.hip: the ray kernel's two passes recompute the same terms bit for bit
NO_CONTRACT = {"traverse.hip", "bp.hip"}
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        flags = list(HIP_FLAGS)
        if src in NO_CONTRACT:
            flags.append("-ffp-contract=off")
        cmd = [hipcc, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    out = os.path.join(HERE, "libmrfmap_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_synth(verbose=False):
    out = os.path.join(HERE, "libmrfmap_synth.so")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "synth.c"), "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_cpp_runner(verbose=False):
    """tests/cpp/run_workload: a caller of the C++ surface (include/mrfmap_hip.hpp)
    that the GPU tests run; linked against libmrfmap_hip.so beside this file."""
    src = os.path.join(ROOT, "tests", "cpp", "run_workload.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "run_workload")
    tmp = out + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), src, "-o", tmp, "-L", HERE,
           "-lmrfmap_hip", "-Wl,-rpath,$ORIGIN/../../mrfmap_amd"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    outs = [build_synth(verbose), build_hip(verbose)]
    runner = build_cpp_runner(verbose)
    return outs + ([runner] if runner else [])


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print(p)
