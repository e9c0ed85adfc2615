"""Synthetic keyframe workloads (libmrfmap_synth.so; host C, deterministic).

Workloads (sizes and seeds are part of each result):
  small     4x4x3 m room, 12 boxes, 4 spheres, 4 cm voxels, 4 keyframes of
            160x120 (f=120, c=(80,60)), seed 1, 5 BP iterations.
  real      same generator at 640x480 (f=525, c=(319.5,239.5)), 13 keyframes,
            1 cm voxels, seed 2, iterate to max marginal change < 1e-4.
  batched   10x10x3 m room, 60 boxes, 10 m range, 64 keyframes drawn from
            {320x240, 640x480, 848x480}, 1 cm voxels, seed 3.
Depth is an exact planar-z ray cast plus N(0, (0.0012 + 0.0019 (z-0.4)^2)^2)
noise and 2 % dropout to 0; poses sit on a ring looking at the room centre
with +-5 deg yaw jitter.  The sensor model of the BP is the reference's
default shape sigma(z) = 1.5 res (GVDBParams.h:48), bias 0.
"""
import ctypes
import os

import numpy as np

from .params import gvdb_params
from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrfmap_synth.so")


class SynthCfg(ctypes.Structure):
    _fields_ = [
        ("room", ctypes.c_float * 3), ("num_boxes", ctypes.c_int), ("box_min", ctypes.c_float),
        ("box_max", ctypes.c_float), ("num_spheres", ctypes.c_int), ("sph_min", ctypes.c_float),
        ("sph_max", ctypes.c_float), ("ring_radius", ctypes.c_float), ("ring_height", ctypes.c_float),
        ("yaw_jitter_deg", ctypes.c_float), ("max_range", ctypes.c_float), ("seed", ctypes.c_uint64),
    ]


_lib_synth = None


def _load():
    global _lib_synth
    if _lib_synth is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m mrfmap_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.mrf_synth_default_room.argtypes = [ctypes.POINTER(SynthCfg), ctypes.c_uint64]
        L.mrf_synth_large_room.argtypes = [ctypes.POINTER(SynthCfg), ctypes.c_uint64]
        L.mrf_synth_create.argtypes = [ctypes.POINTER(SynthCfg), ctypes.POINTER(P)]
        L.mrf_synth_destroy.argtypes = [P]
        L.mrf_synth_ring_pose.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
        L.mrf_synth_render.argtypes = [P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, P]
        L.mrf_synth_occupancy.argtypes = [P, ctypes.c_float, P, ctypes.c_int64, P]
        L.mrf_synth_num_objects.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib_synth = L
    return _lib_synth


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Scene:
    def __init__(self, kind="room", seed=1, **overrides):
        L = _load()
        cfg = SynthCfg()
        (L.mrf_synth_large_room if kind == "large" else L.mrf_synth_default_room)(ctypes.byref(cfg), seed)
        for k, v in overrides.items():
            if k == "room":
                cfg.room[:] = list(v)
            else:
                setattr(cfg, k, v)
        self.cfg = cfg
        self._h = ctypes.c_void_p()
        if L.mrf_synth_create(ctypes.byref(cfg), ctypes.byref(self._h)) != 0:
            raise ValueError("mrf_synth_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            _load().mrf_synth_destroy(self._h)
            self._h = None

    def ring_pose(self, k, n):
        pose = np.zeros((4, 4), dtype=np.float32)
        if _load().mrf_synth_ring_pose(self._h, k, n, _ptr(pose)) != 0:
            raise ValueError("ring_pose")
        return pose

    def render(self, pose, K, rows, cols, noise=True, dropout=0.02, noise_seed=0):
        K = np.asarray(K, dtype=np.float64).reshape(3, 3)
        pose = np.ascontiguousarray(pose, dtype=np.float32)
        depth = np.zeros((rows, cols), dtype=np.float32)
        rc = _load().mrf_synth_render(self._h, _ptr(pose), K[0, 0], K[1, 1], K[0, 2], K[1, 2], rows, cols,
                                      int(bool(noise)), float(dropout), int(noise_seed), _ptr(depth))
        if rc != 0:
            raise ValueError("render")
        return depth

    def occupancy(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        occ = np.zeros(len(pts), dtype=np.uint8)
        _load().mrf_synth_occupancy(self._h, 0.0, _ptr(pts), len(pts), _ptr(occ))
        return occ

    def num_objects(self):
        b, s = ctypes.c_int(), ctypes.c_int()
        _load().mrf_synth_num_objects(self._h, ctypes.byref(b), ctypes.byref(s))
        return b.value, s.value


def intrinsics(rows, cols):
    """Pinhole intrinsics of the three frame shapes of the workloads."""
    table = {(120, 160): (120.0, 80.0, 60.0), (240, 320): (262.5, 159.5, 119.5),
             (480, 640): (525.0, 319.5, 239.5), (480, 848): (423.0, 423.5, 239.5)}
    f, cx, cy = table[(rows, cols)]
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1]], dtype=np.float32)


WORKLOADS = {
    "small": dict(kind="room", seed=1, frames=4, shapes=[(120, 160)], res=0.04, dims=(4.0, 4.0, 3.0), iters=5),
    "real": dict(kind="room", seed=2, frames=13, shapes=[(480, 640)], res=0.01, dims=(4.0, 4.0, 3.0), iters=0),
    "batched": dict(kind="large", seed=3, frames=64, shapes=[(240, 320), (480, 640), (480, 848)], res=0.01,
                    dims=(10.0, 10.0, 3.0), iters=0),
}


def make_workload(name, frames=None, res=None, shapes=None, noise=True, activation=_lib.ACT_FULL_RAY,
                  msg_dtype=_lib.MSG_F32, seed=None, prior=0.1):
    """Return (params, [(pose 4x4, K 3x3, depth)...]) for a named workload."""
    w = dict(WORKLOADS[name])
    if frames is not None:
        w["frames"] = frames
    if res is not None:
        w["res"] = res
    if shapes is not None:
        w["shapes"] = shapes
    if seed is not None:
        w["seed"] = seed
    scene = Scene(w["kind"], w["seed"])
    p = gvdb_params()
    rows, cols = w["shapes"][0]
    p.rows, p.cols = rows, cols
    p.K = intrinsics(rows, cols)
    p.res = float(w["res"])
    p.prior = float(prior)
    p.dims = np.asarray(w["dims"], dtype=np.float32)
    p.lookup_n = 3
    p.const_sigmasq_poly = [1.5 * p.res, 0.0, 0.0]
    p.const_bias_poly = [0.0, 0.0, 0.0]
    p.max_iters = w["iters"] if w["iters"] > 0 else 50
    p.activation = activation
    p.msg_dtype = msg_dtype
    frames = []
    n = w["frames"]
    rng = np.random.default_rng(w["seed"] + 7)
    for k in range(n):
        shp = w["shapes"][0] if len(w["shapes"]) == 1 else w["shapes"][int(rng.integers(len(w["shapes"])))]
        K = intrinsics(*shp)
        pose = scene.ring_pose(k, n)
        depth = scene.render(pose, K, shp[0], shp[1], noise=noise, noise_seed=w["seed"] * 1000 + k)
        frames.append((pose, K, depth))
    return p, frames, scene
