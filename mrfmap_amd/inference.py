"""GVDBInference: Python surface of the BP inference path.

Mirrors the reference's pybind11 class (src/GVDBPyModules.cpp:64-83) and the
C++ class it wraps (include/mrfmap/GVDBInference.h:20-103) for the hot path:
``add_camera``, ``add_camera_with_depth``, ``set_pose``, ``perform_inference``,
``perform_inference_ids``, ``get_occupied_bricks``, ``get_indices_and_alphas``.
Every call goes through the C ABI of libmrfmap_hip.so; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .params import global_params


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class GVDBImage:
    """A depth keyframe (include/mrfmap/GVDBImage.h).  The reference uploads
    it to the GPU on construction; here it is kept on the host until the
    image is attached to a camera."""

    def __init__(self, image, params=None):
        self.params = params or global_params()
        self.depth = np.ascontiguousarray(np.asarray(image, dtype=np.float32))
        if self.depth.shape != (self.params.rows, self.params.cols):
            raise ValueError(f"depth image shape {self.depth.shape} != ({self.params.rows}, {self.params.cols})")

    def get_reproj_points(self):
        """Camera-frame points of valid pixels (GVDBImage.cpp:44-60, with fy for
        the vertical axis and without the reference's dropped last point)."""
        K = np.asarray(self.params.K, dtype=np.float64)
        v, u = np.nonzero(self.depth > 0)
        d = self.depth[v, u].astype(np.float64)
        return np.stack([(u - K[0, 2]) / K[0, 0] * d, (v - K[1, 2]) / K[1, 1] * d, d], axis=1).astype(np.float32)


class GVDBInference:
    def __init__(self, create_context=True, use_gl=False, params=None):
        self.params = (params or global_params()).copy()
        self._L = _lib.load()
        self._cp = self.params.to_c()
        self._h = ctypes.c_void_p()
        _lib.check(self._L.mrf_create(ctypes.byref(self._cp), ctypes.byref(self._h)))
        self.last_stats = None
        self._cam_shapes = {}

    def close(self):
        """Destroy the context.  Its large device blocks stay with the library
        for the next context; ``mrfmap_amd.release_device_cache()`` gives them
        back to the HIP runtime."""
        if getattr(self, "_h", None):
            self._L.mrf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ keyframes
    def add_camera_with_depth(self, pose, depth, K=None):
        pose = np.ascontiguousarray(np.asarray(pose, dtype=np.float32).reshape(4, 4))
        depth = np.ascontiguousarray(np.asarray(depth, dtype=np.float32))
        cid = ctypes.c_uint32()
        if K is None:
            rc = self._L.mrf_add_camera_with_depth(self._h, _ptr(pose), _ptr(depth), depth.shape[0],
                                                   depth.shape[1], _lib.ROW_MAJOR, ctypes.byref(cid))
        else:
            Kc = np.ascontiguousarray(np.asarray(K, dtype=np.float32).reshape(3, 3))
            rc = self._L.mrf_add_camera_with_depth_k(self._h, _ptr(pose), _ptr(Kc), _ptr(depth), depth.shape[0],
                                                     depth.shape[1], _lib.ROW_MAJOR, ctypes.byref(cid))
        _lib.check(rc, self._h)
        self._cam_shapes[cid.value] = depth.shape
        return cid.value

    def add_camera(self, pose, image):
        return self.add_camera_with_depth(pose, image.depth)

    def set_pose(self, cam_id, pose):
        pose = np.ascontiguousarray(np.asarray(pose, dtype=np.float32).reshape(4, 4))
        _lib.check(self._L.mrf_set_pose(self._h, int(cam_id), _ptr(pose), _lib.ROW_MAJOR), self._h)

    def num_cameras(self):
        n = ctypes.c_uint32()
        _lib.check(self._L.mrf_num_cameras(self._h, ctypes.byref(n)), self._h)
        return n.value

    # ------------------------------------------------------------ inference
    def perform_inference(self, max_iters=0, tol=0.0):
        return self.perform_inference_ids(None, False, max_iters=max_iters, tol=tol)

    def perform_inference_ids(self, cam_ids, dry_run=False, max_iters=0, tol=0.0):
        st = _lib.MrfStats()
        if cam_ids is None:
            rc = self._L.mrf_perform_inference(self._h, None, 0, int(max_iters), float(tol), ctypes.byref(st))
        else:
            ids = np.ascontiguousarray(cam_ids, dtype=np.uint32)
            rc = self._L.mrf_perform_inference(self._h, _ptr(ids), len(ids), int(max_iters), float(tol),
                                               ctypes.byref(st))
        _lib.check(rc, self._h)
        self.last_stats = st.as_dict()
        return self.last_stats

    def perform_inference_warm(self, params, max_iters=0, tol=0.0, growth=2.0, warm_after=1):
        """perform_inference with the learned warm start after ``warm_after``
        plain iterations (``params``: MLP_PARAMS floats trained on features
        taken after as many iterations, see mrfmap_amd.warmstart); the warm
        start is undone when a residual exceeds ``growth`` x the previous one
        and 0.05 (``growth`` = 0: never)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (_lib.MLP_PARAMS,):
            raise ValueError(f"params must hold {_lib.MLP_PARAMS} floats")
        st = _lib.MrfStats()
        rc = self._L.mrf_perform_inference_warm(self._h, _ptr(p), int(warm_after), int(max_iters), float(tol),
                                                float(growth), ctypes.byref(st))
        _lib.check(rc, self._h)
        self.last_stats = st.as_dict()
        return self.last_stats

    def set_reference(self, alphas, stop_at=0.0):
        """Reference marginals (num_bricks x 512) for convergence measurements:
        every later perform_inference reports max |alpha - ref| after its last
        iteration in ``ref_error`` and, with ``stop_at`` > 0, stops once it is
        <= stop_at.  ``None`` removes it."""
        if alphas is None:
            _lib.check(self._L.mrf_set_reference(self._h, None, 0, 0, 0.0), self._h)
            return
        a = np.ascontiguousarray(alphas, dtype=np.float32)
        if a.ndim != 2:
            raise ValueError("reference must be num_bricks x 512")
        _lib.check(self._L.mrf_set_reference(self._h, _ptr(a), a.shape[0], a.shape[1], float(stop_at)), self._h)

    def get_trace(self):
        """(residual, reference error) per iteration of the last call."""
        n = ctypes.c_int()
        _lib.check(self._L.mrf_get_trace(self._h, None, None, 0, ctypes.byref(n)), self._h)
        r = np.zeros(max(n.value, 1))
        e = np.zeros(max(n.value, 1))
        _lib.check(self._L.mrf_get_trace(self._h, _ptr(r), _ptr(e), n.value, ctypes.byref(n)), self._h)
        return r[:n.value], e[:n.value]

    def features(self):
        """(X n x 32 float32, pool index n uint32) of the warm-start features of
        the current state (call after perform_inference)."""
        n = ctypes.c_int64()
        _lib.check(self._L.mrf_features(self._h, None, None, 0, ctypes.byref(n)), self._h)
        X = np.zeros((max(n.value, 1), _lib.MLP_FEATURES), dtype=np.float32)
        vox = np.zeros(max(n.value, 1), dtype=np.uint32)
        _lib.check(self._L.mrf_features(self._h, _ptr(X), _ptr(vox), n.value, ctypes.byref(n)), self._h)
        return X[:n.value], vox[:n.value]

    def warm_predictions(self):
        """(pool index, predicted log-odds) of the last warm start."""
        n = ctypes.c_int64()
        _lib.check(self._L.mrf_warm_predictions(self._h, None, None, 0, ctypes.byref(n)), self._h)
        vox = np.zeros(max(n.value, 1), dtype=np.uint32)
        pred = np.zeros(max(n.value, 1), dtype=np.float32)
        _lib.check(self._L.mrf_warm_predictions(self._h, _ptr(vox), _ptr(pred), n.value, ctypes.byref(n)), self._h)
        return vox[:n.value], pred[:n.value]

    # ------------------------------------------------------------ warm-start MLP on this context's GPU
    def mlp_forward(self, params, X):
        p = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, _lib.MLP_FEATURES)
        y = np.zeros(len(X), dtype=np.float32)
        _lib.check(self._L.mrf_mlp_forward(self._h, _ptr(p), _ptr(X), len(X), _ptr(y)), self._h)
        return y

    def mlp_grad(self, params, X, t):
        """(loss, gradient) of 1/(2n) sum (y - t)^2."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, _lib.MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(_lib.MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _lib.check(self._L.mrf_mlp_grad(self._h, _ptr(p), _ptr(X), _ptr(t), len(X), _ptr(g), ctypes.byref(loss)),
                   self._h)
        return loss.value, g

    def mlp_train(self, params, X, t, **opts):
        """Adam training on the GPU; returns (params, per-epoch mean loss).
        opts: epochs, batch, lr, beta1, beta2, eps, lr_decay, seed."""
        o = _lib.MrfTrainOpts()
        self._L.mrf_train_opts_default(ctypes.byref(o))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown training option {k}")
            setattr(o, k, v)
        p = np.array(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, _lib.MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(max(o.epochs, 1), dtype=np.float64)
        _lib.check(self._L.mrf_mlp_train(self._h, _ptr(p), _ptr(X), _ptr(t), len(X), ctypes.byref(o),
                                         _ptr(losses)), self._h)
        return p, losses[:o.epochs]

    def prepare(self):
        _lib.check(self._L.mrf_prepare(self._h), self._h)

    # ------------------------------------------------------------ readback
    def num_bricks(self):
        n = ctypes.c_int64()
        _lib.check(self._L.mrf_num_bricks(self._h, ctypes.byref(n)), self._h)
        return n.value

    def get_occupied_bricks(self, lev=0):
        if lev != 0:
            raise ValueError("only level 0 (8^3 bricks) exists in the brick grid")
        n = self.num_bricks()
        out = np.zeros((n, 3), dtype=np.float32)
        _lib.check(self._L.mrf_get_occupied_bricks(self._h, _ptr(out), n), self._h)
        return out

    def get_indices_and_alphas(self, out):
        """Fill ``out`` (num_bricks x 512 float32, C-contiguous) with the marginals."""
        if not (isinstance(out, np.ndarray) and out.dtype == np.float32 and out.flags.c_contiguous and out.ndim == 2):
            raise TypeError("out must be a C-contiguous float32 2-D array")
        _lib.check(self._L.mrf_get_alphas(self._h, _ptr(out), out.shape[0], out.shape[1]), self._h)
        return out

    def get_alphas(self):
        out = np.zeros((self.num_bricks(), 512), dtype=np.float32)
        return self.get_indices_and_alphas(out)

    def get_beliefs(self):
        out = np.zeros((self.num_bricks(), 512), dtype=np.float32)
        _lib.check(self._L.mrf_get_beliefs(self._h, _ptr(out), out.shape[0], out.shape[1]), self._h)
        return out

    def get_cum_lengths(self, cam_id):
        """Camera cam_id's summed chord length C per voxel (num_bricks x 512, 0 where its rays do not pass)."""
        out = np.zeros((self.num_bricks(), 512), dtype=np.float32)
        _lib.check(self._L.mrf_get_cum_lengths(self._h, int(cam_id), _ptr(out), out.shape[0], out.shape[1]),
                   self._h)
        return out

    def get_traversal(self, cam_id):
        nr, nc = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._L.mrf_get_traversal(self._h, int(cam_id), ctypes.byref(nr), ctypes.byref(nc), None, None),
                   self._h)
        rays = np.zeros((nr.value, 6), dtype=np.uint32)
        recs = np.zeros((nc.value, 2), dtype=np.uint32)
        _lib.check(self._L.mrf_get_traversal(self._h, int(cam_id), None, None, _ptr(rays), _ptr(recs)), self._h)
        return rays, recs

    # ------------------------------------------------------------ map evaluation
    def _cam_shape(self, cam_id):
        return self._cam_shapes.get(int(cam_id), (self.params.rows, self.params.cols))

    def evaluate(self, cam_id=None, pose=None, depth=None, K=None, shape=None):
        """Per-pixel log-likelihood, expected depth, accuracy and simple
        accuracy images of the current map seen from camera ``cam_id`` or from
        ``pose`` with ``depth`` (row-major), and the image-level reductions."""
        if cam_id is not None:
            rows, cols = shape or self._cam_shape(cam_id)
        else:
            rows, cols = (np.asarray(depth).shape if depth is not None else (shape or (self.params.rows, self.params.cols)))
        imgs = [np.zeros((rows, cols), dtype=np.float32) for _ in range(4)]
        ev = _lib.MrfEval()
        if cam_id is not None:
            rc = self._L.mrf_evaluate_camera(self._h, int(cam_id), *[_ptr(i) for i in imgs], ctypes.byref(ev))
        else:
            pose = np.ascontiguousarray(np.asarray(pose, dtype=np.float32).reshape(4, 4))
            d = None if depth is None else np.ascontiguousarray(np.asarray(depth, dtype=np.float32))
            k = None if K is None else np.ascontiguousarray(np.asarray(K, dtype=np.float32).reshape(3, 3))
            rc = self._L.mrf_evaluate_at_pose(self._h, _ptr(pose), None if k is None else _ptr(k),
                                              None if d is None else _ptr(d), rows, cols, _lib.ROW_MAJOR,
                                              *[_ptr(i) for i in imgs], ctypes.byref(ev))
        _lib.check(rc, self._h)
        out = ev.as_dict()
        out.update(loglik=imgs[0], expected_depth=imgs[1], accuracy_image=imgs[2], simple_accuracy_image=imgs[3])
        return out

    def compute_likelihood(self, cam_id):
        """Sum of the per-pixel log-likelihoods of camera ``cam_id``'s depth
        image under the current map (GVDBInference.cpp:180-186)."""
        return self.evaluate(cam_id)["log_likelihood"]

    def get_likelihood_image(self, cam_id, out):
        r = self.evaluate(cam_id, shape=out.shape)
        out[...] = r["loglik"]
        return r["log_likelihood"]

    def get_likelihood_image_at_pose(self, pose, image, out):
        depth = image.depth if isinstance(image, GVDBImage) else image
        r = self.evaluate(pose=pose, depth=depth)
        out[...] = r["loglik"]
        return r["log_likelihood"]

    def get_expected_depth_image(self, cam_id, out):
        out[...] = self.evaluate(cam_id, shape=out.shape)["expected_depth"]

    def get_accuracy_image_at_pose(self, pose, image, out):
        depth = image.depth if isinstance(image, GVDBImage) else image
        r = self.evaluate(pose=pose, depth=depth)
        out[...] = r["accuracy_image"]
        return r["accuracy"]

    def get_simple_accuracy_at_pose(self, pose, image):
        """(accuracy, tp, fp, tn, fn) of the first-voxel-above-acc_thresh test
        (compute_tp_fp_tn_fn, gvdbhelper.cu:40-63)."""
        depth = image.depth if isinstance(image, GVDBImage) else image
        r = self.evaluate(pose=pose, depth=depth)
        tn, fn, tp, fp = r["simple_counts"]
        return r["simple_accuracy"], tp, fp, tn, fn

    # ------------------------------------------------------------ external maps
    def load_map(self, points, occupancies):
        """Load an occupancy map: points (n x 3, voxel units) and their
        occupancy probabilities (GVDBMapLikelihoodEstimator::load_map)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        occ = np.ascontiguousarray(np.asarray(occupancies, dtype=np.float32).reshape(-1))
        if len(occ) != len(pts):
            raise ValueError("one occupancy per point")
        _lib.check(self._L.mrf_load_map(self._h, _ptr(pts), _ptr(occ), len(pts)), self._h)
        return True

    def set_launch_mode(self, mode):
        """_lib.LAUNCH_GRAPH (iterations replayed from a captured HIP graph, the
        default) or _lib.LAUNCH_TIMED (kernels launched one by one with
        per-kernel device timings in the returned statistics)."""
        _lib.check(self._L.mrf_set_launch_mode(self._h, int(mode)), self._h)

    def set_acc_thresh(self, thresh):
        self.params.acc_thresh = float(thresh)
        _lib.check(self._L.mrf_set_eval_thresholds(self._h, float(self.params.occ_thresh), float(thresh)), self._h)


class GVDBMapLikelihoodEstimator(GVDBInference):
    """Scores an occupancy map against depth keyframes
    (include/mrfmap/GVDBMapLikelihoodEstimator.h, GVDBPyModules.cpp:85-105):
    ``load_map`` an external map (e.g. an OctoMap's occupied leaves), then
    the likelihood / accuracy images at any pose."""

    def __init__(self, params=None, use_gl=False):
        super().__init__(True, use_gl, params=params)

    def activate_camera(self, cam_id):
        return int(cam_id)

    def compute_likelihood(self, cam_id):
        return super().compute_likelihood(cam_id)

    def get_rendered_image_at_pose(self, pose, image, out):
        """Expected depth image of the loaded map from ``pose``."""
        depth = image.depth if isinstance(image, GVDBImage) else image
        out[...] = self.evaluate(pose=pose, depth=depth)["expected_depth"]
