"""CPU oracle of the MRFMap BP inference path — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker / CPU baseline.  The
product (``mrfmap_amd``) never imports it.

It wraps ``liboracle.so`` (oracle.c, octomap_eq.c; scalar C, FP64), a
restatement of the reference's live BP math:
  traversal  GVDB HDDA driven by src/cuda/gvdbkernels.cu:519-571, 689-1038,
             1164-1226, 1276-1302 (GVDB not vendored: rule defined here)
  messages   src/cuda/gvdbkernels.cu:46-72, 689-1038, 1196-1208
  epilogue   src/cuda/gvdbkernels.cu:457-479
  schedule   src/GVDBInference.cpp:100-178
  readback   src/GVDBMRFMap.cpp:137-208, src/cuda/gvdbkernels.cu:360-373
  octomap    src/GVDBOctomapWrapper.cpp:3-43 + OctoMap's published insertion

The reference itself (C++/CUDA + GVDB fetched over the network) cannot be
built in this image, so there is no ``oracle/_ref``.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SOURCES = ["oracle.c", "octomap_eq.c"]


def build(verbose=False):
    """Compile liboracle.so with gcc (IEEE double, no FMA contraction)."""
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           *[os.path.join(HERE, s) for s in SOURCES], "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_float * 9), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
        ("res", ctypes.c_float), ("prior", ctypes.c_float), ("dims", ctypes.c_float * 3),
        ("lookup_n", ctypes.c_int), ("sigma_poly", ctypes.c_float * 3), ("bias_poly", ctypes.c_float * 3),
        ("activation", ctypes.c_int), ("mask_848_border", ctypes.c_int),
        ("sigma_lookup", ctypes.c_void_p), ("bias_lookup", ctypes.c_void_p), ("brick_local_vis", ctypes.c_int),
    ]


def _bind(L):
    """Declare the C signatures of liboracle (or of a variant built from a
    modified copy of its sources, see load_variant)."""
    P = ctypes.c_void_p
    f64p = ctypes.POINTER(ctypes.c_double)
    L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(P)]
    L.orc_destroy.argtypes = [P]
    L.orc_add_camera.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int]
    L.orc_set_pose.argtypes = [P, ctypes.c_int, P]
    L.orc_set_beliefs.argtypes = [P, P, P]
    L.orc_load_map.argtypes = [P, P, P, ctypes.c_int64]
    L.orc_evaluate.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, P, P, P]
    L.orc_prepare.argtypes = [P]
    L.orc_num_bricks.argtypes = [P]
    L.orc_num_bricks.restype = ctypes.c_int64
    L.orc_brick_coords.argtypes = [P, P]
    L.orc_get_traversal.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int64), P, P]
    L.orc_perform_inference.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, f64p,
                                        ctypes.POINTER(ctypes.c_int64)]
    L.orc_continue_inference.argtypes = [P, ctypes.c_int, ctypes.c_double, f64p]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_threads.restype = None
    L.orc_get_threads.argtypes = []
    L.orc_get_threads.restype = ctypes.c_int
    L.orc_features.argtypes = [P, P, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.orc_warm_start.argtypes = [P, P, P, ctypes.c_int64]
    L.orc_get_alphas.argtypes = [P, P]
    L.orc_get_beliefs.argtypes = [P, P]
    L.orc_get_cum_lengths.argtypes = [P, ctypes.c_int, P]
    L.orc_ray_messages.argtypes = [ctypes.c_int, P, P, P, P, ctypes.c_double, P, P, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int, P]
    L.orc_ray_messages.restype = ctypes.c_double
    L.orc_ray_messages_ex.argtypes = [ctypes.c_int, P, P, P, P, P, ctypes.c_double, P, P, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, P]
    L.orc_ray_messages_ex.restype = ctypes.c_double
    L.orc_ray_messages_bruteforce.argtypes = [ctypes.c_int, P, P, ctypes.c_double, P]
    L.orc_depth_psi.argtypes = [ctypes.c_double, ctypes.c_double, P, P, ctypes.c_int]
    L.orc_depth_psi.restype = ctypes.c_double
    L.orc_walk_ray.argtypes = [P, P, P, ctypes.c_double, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int,
                               P, P, P, ctypes.POINTER(ctypes.c_int)]
    L.oct_create.argtypes = [ctypes.c_double, P, P, ctypes.c_double, ctypes.POINTER(P)]
    L.oct_destroy.argtypes = [P]
    L.oct_insert_point_cloud.argtypes = [P, P, ctypes.c_int64, P]
    L.oct_insert_point_cloud.restype = ctypes.c_int64
    L.oct_num_nodes.argtypes = [P]
    L.oct_num_nodes.restype = ctypes.c_int64
    L.oct_logodds_at.argtypes = [P, P]
    L.oct_logodds_at.restype = ctypes.c_double
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(ctypes.CDLL(LIB))
    return _lib


def load_variant(so_path):
    """A liboracle built elsewhere (tests build planted-bug variants of
    oracle.c to show that a check catches them)."""
    return _bind(ctypes.CDLL(so_path))


def build_variant(replacements, out_dir, name="variant"):
    """Compile a copy of the oracle with textual replacements [(old, new)]
    (each `old` must occur exactly once in oracle.c) into out_dir and load
    it.  Tests use this to plant a mistake in the oracle, on the CPU, and
    show that a check fails on it."""
    src = open(os.path.join(HERE, "oracle.c")).read()
    for old, new in replacements:
        n = src.count(old)
        if n != 1:
            raise ValueError(f"replacement target occurs {n} times: {old!r}")
        src = src.replace(old, new)
    c_path = os.path.join(out_dir, name + ".c")
    so_path = os.path.join(out_dir, name + ".so")
    with open(c_path, "w") as f:
        f.write(src)
    cmd = ["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-I", HERE,
           c_path, os.path.join(HERE, "octomap_eq.c"), "-o", so_path, "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("variant build failed:\n" + r.stdout)
    return load_variant(so_path)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 scalar BP over the same keyframes the GPU path sees."""

    def __init__(self, K, rows, cols, res, prior, dims, sigma_poly=None, bias_poly=None, lookup_n=3,
                 activation=1, mask_848_border=1, library=None, sigma_lookup=None, bias_lookup=None,
                 brick_local_vis=0):
        """sigma_lookup / bias_lookup: optional per-cell tables of
        ceil(rows/20) * ceil(cols/20) * lookup_n coefficients (then every
        keyframe must be rows x cols).  brick_local_vis: 1 for the
        reference's pre pass, whose visibility restarts at every brick until
        a tau survives FP32 flush-to-zero; 0 carries it along the ray."""
        p = OrcParams()
        p.K[:] = [float(v) for v in np.asarray(K, dtype=np.float32).reshape(9)]
        p.rows, p.cols = int(rows), int(cols)
        p.res, p.prior = float(res), float(prior)
        p.dims[:] = [float(v) for v in dims]
        p.lookup_n = int(lookup_n)
        sp = sigma_poly if sigma_poly is not None else [1.5 * res, 0.0, 0.0]
        bp = bias_poly if bias_poly is not None else [0.0, 0.0, 0.0]
        p.sigma_poly[:] = [float(v) for v in list(sp) + [0.0] * (3 - len(sp))]
        p.bias_poly[:] = [float(v) for v in list(bp) + [0.0] * (3 - len(bp))]
        p.activation = int(activation)
        p.mask_848_border = int(mask_848_border)
        p.brick_local_vis = int(brick_local_vis)
        if sigma_lookup is not None and bias_lookup is not None:
            tab_sig = np.ascontiguousarray(sigma_lookup, dtype=np.float32).ravel()
            tab_bia = np.ascontiguousarray(bias_lookup, dtype=np.float32).ravel()
            need = -(-int(rows) // 20) * -(-int(cols) // 20) * int(lookup_n)
            if tab_sig.size < need or tab_bia.size < need:
                raise ValueError(f"lookup tables need {need} coefficients")
            p.sigma_lookup, p.bias_lookup = tab_sig.ctypes.data, tab_bia.ctypes.data
        self._L = library if library is not None else lib()
        self._h = ctypes.c_void_p()
        rc = self._L.orc_create(ctypes.byref(p), ctypes.byref(self._h))
        if rc != 0:
            raise ValueError(f"orc_create failed ({rc})")
        self.K = np.asarray(K, dtype=np.float32).reshape(9)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_destroy(h)
            self._h = None

    def add_camera(self, pose, depth, K=None):
        pose = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        Kc = self.K if K is None else np.ascontiguousarray(K, dtype=np.float32).reshape(9)
        rc = self._L.orc_add_camera(self._h, _ptr(pose), _ptr(Kc), _ptr(depth), depth.shape[0], depth.shape[1])
        if rc < 0:
            raise ValueError(f"orc_add_camera failed ({rc})")
        return rc

    def set_pose(self, cam, pose):
        pose = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
        if self._L.orc_set_pose(self._h, int(cam), _ptr(pose)) != 0:
            raise IndexError(cam)

    def prepare(self):
        self._L.orc_prepare(self._h)

    def load_map(self, points, occupancies):
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        occ = np.ascontiguousarray(occupancies, dtype=np.float32).reshape(-1)
        if self._L.orc_load_map(self._h, _ptr(pts), _ptr(occ), len(pts)) != 0:
            raise RuntimeError("orc_load_map failed")

    def set_beliefs(self, B, upd):
        """Test hook: beliefs (num_bricks x 512 log-odds) and update flags."""
        B = np.ascontiguousarray(B, dtype=np.float64)
        upd = np.ascontiguousarray(upd, dtype=np.uint8)
        if self._L.orc_set_beliefs(self._h, _ptr(B), _ptr(upd)) != 0:
            raise RuntimeError("orc_set_beliefs failed")

    def evaluate(self, pose, depth, K=None, acc_thresh=0.5, shape=None):
        """Map evaluation images and their reductions (sum of log-likelihoods,
        accuracy = #1 / (#pixels - #-1), simple accuracy class counts)."""
        pose = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
        Kc = self.K if K is None else np.ascontiguousarray(K, dtype=np.float32).reshape(9)
        d = None if depth is None else np.ascontiguousarray(depth, dtype=np.float32)
        rows, cols = d.shape if d is not None else shape
        imgs = [np.zeros((rows, cols), dtype=np.float64) for _ in range(4)]
        if self._L.orc_evaluate(self._h, _ptr(pose), _ptr(Kc), None if d is None else _ptr(d), rows, cols,
                              float(acc_thresh), *[_ptr(i) for i in imgs]) != 0:
            raise RuntimeError("orc_evaluate failed")
        ll, ed, acc, sacc = imgs
        n = rows * cols
        acc_counts = [int((acc == k).sum()) for k in (-1, 0, 1)]
        simple_counts = [int((sacc == k).sum()) for k in (-1, 0, 1, 2)]
        return {"loglik": ll, "expected_depth": ed, "accuracy_image": acc, "simple_accuracy_image": sacc,
                "log_likelihood": float(ll.sum()), "accuracy_counts": acc_counts, "simple_counts": simple_counts,
                "accuracy": acc_counts[2] / (n - acc_counts[0]) if n > acc_counts[0] else 0.0,
                "simple_accuracy": (simple_counts[2] + simple_counts[0]) / n}

    def num_bricks(self):
        return int(self._L.orc_num_bricks(self._h))

    def brick_coords(self):
        out = np.zeros((self.num_bricks(), 3), dtype=np.int32)
        self._L.orc_brick_coords(self._h, _ptr(out))
        return out

    def traversal(self, cam):
        nr, nc = ctypes.c_int64(), ctypes.c_int64()
        self._L.orc_get_traversal(self._h, int(cam), ctypes.byref(nr), ctypes.byref(nc), None, None)
        rays = np.zeros((nr.value, 6), dtype=np.uint32)
        recs = np.zeros((nc.value, 2), dtype=np.uint32)
        self._L.orc_get_traversal(self._h, int(cam), None, None, _ptr(rays), _ptr(recs))
        return rays, recs

    def perform_inference(self, max_iters, tol=0.0, ids=None):
        res = ctypes.c_double()
        upd = ctypes.c_int64()
        if ids is not None:
            ida = np.ascontiguousarray(ids, dtype=np.int32)
            it = self._L.orc_perform_inference(self._h, _ptr(ida), len(ida), int(max_iters), float(tol),
                                             ctypes.byref(res), ctypes.byref(upd))
        else:
            it = self._L.orc_perform_inference(self._h, None, 0, int(max_iters), float(tol), ctypes.byref(res),
                                             ctypes.byref(upd))
        if it < 0:
            raise RuntimeError(f"orc_perform_inference failed ({it})")
        return {"iters": it, "residual": res.value, "updates_per_iter": upd.value}

    def continue_inference(self, max_iters, tol=0.0):
        """Iterations without the per-call message reset."""
        res = ctypes.c_double()
        it = self._L.orc_continue_inference(self._h, int(max_iters), float(tol), ctypes.byref(res))
        if it < 0:
            raise RuntimeError(f"orc_continue_inference failed ({it})")
        return {"iters": it, "residual": res.value}

    def features(self):
        """(X n x 32 float32, pool index n) of the warm-start features."""
        n = ctypes.c_int64()
        self._L.orc_features(self._h, None, None, 0, ctypes.byref(n))
        X = np.zeros((max(n.value, 1), 32), dtype=np.float32)
        vox = np.zeros(max(n.value, 1), dtype=np.uint32)
        rc = self._L.orc_features(self._h, _ptr(X), _ptr(vox), n.value, ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"orc_features failed ({rc})")
        return X[:n.value], vox[:n.value]

    def warm_start(self, vox, pred):
        vox = np.ascontiguousarray(vox, dtype=np.uint32)
        pred = _f64(pred)
        if self._L.orc_warm_start(self._h, _ptr(vox), _ptr(pred), len(vox)) != 0:
            raise RuntimeError("orc_warm_start failed")

    def alphas(self):
        out = np.zeros((self.num_bricks(), 512), dtype=np.float64)
        self._L.orc_get_alphas(self._h, _ptr(out))
        return out

    def beliefs(self):
        out = np.zeros((self.num_bricks(), 512), dtype=np.float64)
        self._L.orc_get_beliefs(self._h, _ptr(out))
        return out

    def cum_lengths(self, cam):
        """Camera cam's FP64 summed chord length C per voxel (0 where its rays do not pass)."""
        out = np.zeros((self.num_bricks(), 512), dtype=np.float64)
        if self._L.orc_get_cum_lengths(self._h, int(cam), _ptr(out)) != 0:
            raise RuntimeError("orc_get_cum_lengths failed")
        return out


def set_threads(n):
    """Threads of the oracle's BP loop (0: OpenMP's default)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def ray_messages(nin, length, s, d, sigma_poly, bias_poly=(0.0, 0.0, 0.0), live=None, lookup_n=3,
                 ran_out=False, psi_inf=1.0, mode=0, bstart=None):
    """Outgoing messages p_i of one ray; returns (p, T).  bstart (0/1 per
    record): brick entries for the reference's pre-pass visibility restart."""
    nin, length, s = _f64(nin), _f64(length), _f64(s)
    n = len(nin)
    live = np.ones(n, dtype=np.int32) if live is None else np.ascontiguousarray(live, dtype=np.int32)
    sg = _f64(list(sigma_poly) + [0.0] * (3 - len(sigma_poly)))
    bi = _f64(list(bias_poly) + [0.0] * (3 - len(bias_poly)))
    p = np.zeros(n, dtype=np.float64)
    bs = None if bstart is None else np.ascontiguousarray(bstart, dtype=np.int32)
    T = lib().orc_ray_messages_ex(n, _ptr(nin), _ptr(length), _ptr(s), _ptr(live), None if bs is None else _ptr(bs),
                                  float(d), _ptr(sg), _ptr(bi), int(lookup_n), int(bool(ran_out)), float(psi_inf),
                                  int(mode), _ptr(p))
    return p, T


def ray_messages_bruteforce(nin, psi, psi_inf):
    nin, psi = _f64(nin), _f64(psi)
    p = np.zeros(len(nin), dtype=np.float64)
    if lib().orc_ray_messages_bruteforce(len(nin), _ptr(nin), _ptr(psi), float(psi_inf), _ptr(p)) != 0:
        raise ValueError("n out of range")
    return p


def depth_psi(s, d, sigma_poly, bias_poly=(0.0, 0.0, 0.0), lookup_n=3):
    sg = _f64(list(sigma_poly) + [0.0] * (3 - len(sigma_poly)))
    bi = _f64(list(bias_poly) + [0.0] * (3 - len(bias_poly)))
    return lib().orc_depth_psi(float(s), float(d), _ptr(sg), _ptr(bi), int(lookup_n))


def walk_ray(O, D, N, depth=0.0, trunc=False, sigma_poly=(0.015, 0, 0), bias_poly=(0, 0, 0), lookup_n=3, cap=100000):
    O, D = _f64(O), _f64(D)
    Na = np.ascontiguousarray(N, dtype=np.int32)
    sg, bi = _f64(sigma_poly), _f64(bias_poly)
    vox = np.zeros((cap, 3), dtype=np.int32)
    ln = np.zeros(cap, dtype=np.float64)
    zx = np.zeros(cap, dtype=np.float64)
    ro = ctypes.c_int()
    m = lib().orc_walk_ray(_ptr(O), _ptr(D), _ptr(Na), float(depth), int(bool(trunc)), _ptr(sg), _ptr(bi),
                           int(lookup_n), cap, _ptr(vox), _ptr(ln), _ptr(zx), ctypes.byref(ro))
    return vox[:m], ln[:m], zx[:m], bool(ro.value)


class OctomapEq:
    """OctoMap-equivalent insertion (CPU comparator; not OctoMap itself)."""

    def __init__(self, res, dims, occ_thresh=0.5):
        bmin = _f64([-dims[0] * 0.5, -dims[1] * 0.5, 0.0])
        bmax = _f64([dims[0] * 0.5, dims[1] * 0.5, dims[2]])
        self._h = ctypes.c_void_p()
        if lib().oct_create(float(res), _ptr(bmin), _ptr(bmax), float(occ_thresh), ctypes.byref(self._h)) != 0:
            raise ValueError("oct_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oct_destroy(h)
            self._h = None

    def insert_point_cloud(self, pts, origin):
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        o = _f64(origin)
        return int(lib().oct_insert_point_cloud(self._h, _ptr(pts), len(pts), _ptr(o)))

    def num_nodes(self):
        return int(lib().oct_num_nodes(self._h))

    def logodds_at(self, p):
        return lib().oct_logodds_at(self._h, _ptr(_f64(p)))
