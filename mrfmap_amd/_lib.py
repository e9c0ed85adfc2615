"""ctypes bindings of libmrfmap_hip.so (C ABI in include/mrfmap_hip.h).

The library is built in-tree by ``mrfmap_amd.build``.  There is no CPU
fallback: if the shared library is missing or no HIP device is present the
calls fail loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrfmap_hip.so")

MRF_OK = 0
ERRORS = {-1: "MRF_E_INVALID", -2: "MRF_E_HIP", -3: "MRF_E_NOMEM", -4: "MRF_E_SHAPE", -5: "MRF_E_RANGE",
          -6: "MRF_E_STATE", -7: "MRF_E_LIMIT"}
ROW_MAJOR, COL_MAJOR = 0, 1
ACT_ENDPOINT, ACT_FULL_RAY = 0, 1
MSG_F32, MSG_F16, MSG_BF16, MSG_MIXED_F16, MSG_MIXED_BF16 = 0, 1, 2, 3, 4
VIS_CARRY, VIS_BRICK_LOCAL = 0, 1
# ray-kernel schedules (records loaded ahead in the front loop, in the window passes)
LAUNCH_GRAPH, LAUNCH_TIMED = 0, 1
RAY_SCHEDULES = {0: (8, 4), 1: (8, 8), 2: (16, 4), 3: (16, 8), 4: (4, 4), 5: (16, 16)}

# every symbol include/mrfmap_hip.h declares
EXPORTS = [
    "mrf_params_default", "mrf_create", "mrf_destroy", "mrf_add_camera_with_depth",
    "mrf_add_camera_with_depth_k", "mrf_set_pose", "mrf_num_cameras", "mrf_perform_inference",
    "mrf_num_bricks", "mrf_get_occupied_bricks", "mrf_get_alphas", "mrf_get_beliefs", "mrf_get_cum_lengths",
    "mrf_prepare",
    "mrf_get_traversal", "mrf_last_error", "mrf_evaluate_camera", "mrf_evaluate_at_pose", "mrf_load_map",
    "mrf_set_eval_thresholds",
    "mrf_set_launch_mode", "mrf_train_opts_default", "mrf_perform_inference_warm", "mrf_features",
    "mrf_warm_predictions", "mrf_mlp_forward", "mrf_mlp_grad", "mrf_mlp_train", "mrf_set_reference", "mrf_get_trace",
    "mrf_mlp_grad_noalloc", "mrf_release_device_cache",
]
MLP_FEATURES, MLP_PARAMS = 32, 6337
MLP_NOALLOC_MAX_SAMPLES = 16384


class MrfParams(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_float * 9), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
        ("res", ctypes.c_float), ("prior", ctypes.c_float), ("dims", ctypes.c_float * 3),
        ("occ_thresh", ctypes.c_float), ("acc_thresh", ctypes.c_float), ("max_iters", ctypes.c_int),
        ("lookup_n", ctypes.c_int), ("const_sigmasq_poly", ctypes.c_float * 3),
        ("const_bias_poly", ctypes.c_float * 3), ("depth_sigma_lookup", ctypes.c_void_p),
        ("depth_bias_lookup", ctypes.c_void_p), ("activation", ctypes.c_int), ("msg_dtype", ctypes.c_int),
        ("mask_848_border", ctypes.c_int), ("brick_local_vis", ctypes.c_int),
    ]


class MrfStats(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int), ("converged", ctypes.c_int), ("residual", ctypes.c_double),
        ("updates_per_iter", ctypes.c_int64), ("pairs_per_iter", ctypes.c_int64), ("rays", ctypes.c_int64),
        ("list_records", ctypes.c_int64), ("t_geometry_ms", ctypes.c_double), ("t_iters_ms", ctypes.c_double),
        ("t_ray_kernel_ms", ctypes.c_double), ("t_epilogue_ms", ctypes.c_double),
        ("t_total_ms", ctypes.c_double), ("ray_sched", ctypes.c_int), ("t_autotune_ms", ctypes.c_double),
        ("window_updates_per_iter", ctypes.c_int64), ("graph", ctypes.c_int), ("warm", ctypes.c_int),
        ("warm_voxels", ctypes.c_int64), ("t_warm_ms", ctypes.c_double), ("ref_error", ctypes.c_double),
        ("static_front_updates", ctypes.c_int64),
        ("front_slots", ctypes.c_int64), ("window_slots", ctypes.c_int64),
        ("t_pair_msg_ms", ctypes.c_double), ("t_front_ms", ctypes.c_double), ("t_window_ms", ctypes.c_double),
        ("t_seg_sum_ms", ctypes.c_double), ("t_seg_update_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MrfTrainOpts(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int), ("batch", ctypes.c_int), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("lr_decay", ctypes.c_float), ("seed", ctypes.c_uint64),
    ]


class MrfEval(ctypes.Structure):
    _fields_ = [
        ("log_likelihood", ctypes.c_double), ("accuracy", ctypes.c_double), ("simple_accuracy", ctypes.c_double),
        ("accuracy_counts", ctypes.c_int64 * 3), ("simple_counts", ctypes.c_int64 * 4),
    ]

    def as_dict(self):
        return {"log_likelihood": self.log_likelihood, "accuracy": self.accuracy,
                "simple_accuracy": self.simple_accuracy, "accuracy_counts": list(self.accuracy_counts),
                "simple_counts": list(self.simple_counts)}


class MrfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libmrfmap_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m mrfmap_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32p = ctypes.POINTER(ctypes.c_uint32)
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.mrf_params_default.argtypes = [ctypes.POINTER(MrfParams)]
    L.mrf_params_default.restype = None
    L.mrf_create.argtypes = [ctypes.POINTER(MrfParams), ctypes.POINTER(P)]
    L.mrf_destroy.argtypes = [P]
    L.mrf_destroy.restype = None
    L.mrf_add_camera_with_depth.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p]
    L.mrf_add_camera_with_depth_k.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p]
    L.mrf_set_pose.argtypes = [P, ctypes.c_uint32, P, ctypes.c_int]
    L.mrf_num_cameras.argtypes = [P, u32p]
    L.mrf_perform_inference.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.POINTER(MrfStats)]
    L.mrf_num_bricks.argtypes = [P, i64p]
    L.mrf_get_occupied_bricks.argtypes = [P, P, ctypes.c_int64]
    L.mrf_get_alphas.argtypes = [P, P, ctypes.c_int64, ctypes.c_int64]
    L.mrf_get_beliefs.argtypes = [P, P, ctypes.c_int64, ctypes.c_int64]
    L.mrf_get_cum_lengths.argtypes = [P, ctypes.c_uint32, P, ctypes.c_int64, ctypes.c_int64]
    L.mrf_prepare.argtypes = [P]
    L.mrf_get_traversal.argtypes = [P, ctypes.c_uint32, i64p, i64p, P, P]
    L.mrf_evaluate_camera.argtypes = [P, ctypes.c_uint32, P, P, P, P, ctypes.POINTER(MrfEval)]
    L.mrf_evaluate_at_pose.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, P,
                                       ctypes.POINTER(MrfEval)]
    L.mrf_load_map.argtypes = [P, P, P, ctypes.c_int64]
    L.mrf_set_eval_thresholds.argtypes = [P, ctypes.c_float, ctypes.c_float]
    L.mrf_set_launch_mode.argtypes = [P, ctypes.c_int]
    L.mrf_train_opts_default.argtypes = [ctypes.POINTER(MrfTrainOpts)]
    L.mrf_train_opts_default.restype = None
    L.mrf_perform_inference_warm.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(MrfStats)]
    L.mrf_features.argtypes = [P, P, P, ctypes.c_int64, i64p]
    L.mrf_set_reference.argtypes = [P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_double]
    L.mrf_get_trace.argtypes = [P, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mrf_warm_predictions.argtypes = [P, P, P, ctypes.c_int64, i64p]
    L.mrf_mlp_forward.argtypes = [P, P, P, ctypes.c_int64, P]
    L.mrf_mlp_grad.argtypes = [P, P, P, P, ctypes.c_int64, P, ctypes.POINTER(ctypes.c_double)]
    L.mrf_mlp_train.argtypes = [P, P, P, P, ctypes.c_int64, ctypes.POINTER(MrfTrainOpts), P]
    L.mrf_mlp_grad_noalloc.argtypes = [P, P, P, P, ctypes.c_int64, P, P, P]
    L.mrf_last_error.argtypes = [P]
    L.mrf_last_error.restype = ctypes.c_char_p
    L.mrf_release_device_cache.argtypes = []
    _lib = L
    return L


def release_device_cache():
    """Give the large device blocks the library keeps after freeing them back
    to the HIP runtime (mrf_release_device_cache)."""
    check(load().mrf_release_device_cache())


def check(rc, h=None):
    if rc != MRF_OK:
        msg = ""
        if h:
            m = load().mrf_last_error(h)
            msg = m.decode() if m else ""
        raise MrfError(rc, msg)
    return rc
