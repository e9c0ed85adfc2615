"""gvdb_params: the reference's parameter object (include/mrfmap/GVDBParams.h:11-215).

Field names, defaults, the YAML keys read by ``load_from_file`` and the
process-global instance set by ``set_from_python`` (src/GVDBPyModules.cpp:23-29)
follow the reference.  Config errors raise ``ValueError`` instead of the
reference's ``exit(-1)`` (GVDBParams.h:114,181).
"""
import math
import os

import numpy as np
import yaml

from . import _lib

LOOKUP_PX = 20
LOOKUP_NUM = 3096


class gvdb_params:  # noqa: N801  (reference class name)
    hfov = 0.01

    def __init__(self):
        self.res = 0.01
        self.occ_thresh = 0.1
        self.acc_thresh = 0.5
        self.prior = 0.01
        self.max_iters = 1
        self.dims = np.array([1.0, 1.0, 1.0], dtype=np.float32)
        self.rows = 1
        self.cols = 1
        K = np.eye(3, dtype=np.float32)
        K[0, 0] = K[1, 1] = (self.cols / 2.0) / math.tan(self.hfov / 2.0)
        K[0, 2] = self.cols * 0.5
        K[1, 2] = self.rows * 0.5
        self.K = K
        self.use_polys = False
        self.lookup_n = 3
        self.const_sigmasq_poly = [1.5 * self.res, 0.0, 0.0]
        self.const_bias_poly = [0.0, 0.0, 0.0]
        self.depth_sigma_lookup = None
        self.depth_bias_lookup = None
        self.gvdb_map_config = [1, 3, 3, 3, 3]
        # extensions of this implementation (no reference counterpart)
        self.activation = _lib.ACT_ENDPOINT
        self.msg_dtype = _lib.MSG_F32
        self.mask_848_border = True
        # pre-pass visibility: _lib.VIS_BRICK_LOCAL (default) restarts it at
        # every brick entry until a tau survives FP32 flush-to-zero, as the
        # reference's pre pass does (gvdbkernels.cu:708-715); _lib.VIS_CARRY
        # carries it along the ray, as the post pass does
        self.brick_local_vis = _lib.VIS_BRICK_LOCAL

    def lookup_cells(self):
        return math.ceil(self.cols / LOOKUP_PX) * math.ceil(self.rows / LOOKUP_PX)

    def populate_lookup_tables(self):
        """Constant polynomial in every 20x20-pixel cell (GVDBParams.h:93-104)."""
        n = self.lookup_cells()
        self.depth_sigma_lookup = np.tile(np.asarray(self.const_sigmasq_poly[: self.lookup_n], np.float32), n)
        self.depth_bias_lookup = np.tile(np.asarray(self.const_bias_poly[: self.lookup_n], np.float32), n)

    def load_from_file(self, file_path):
        """Read the ``gvdb_params`` node of a YAML config (GVDBParams.h:106-214)."""
        with open(file_path) as f:
            root = yaml.safe_load(f)
        node = (root or {}).get("gvdb_params")
        if node is None:
            raise ValueError("[gvdb_params] Could not read [gvdb_params]!")
        if "rows" in node:
            self.rows = int(float(node["rows"]))
        if "cols" in node:
            self.cols = int(float(node["cols"]))
        if "dims" in node:
            self.dims = np.asarray([float(v) for v in node["dims"][:3]], dtype=np.float32)
        if "res" in node:
            self.res = float(node["res"])
        if "use_polys" not in node:
            raise ValueError("[gvdb_params] Please configure use_polys in config file.")
        self.use_polys = bool(node["use_polys"])
        if self.use_polys:
            self.lookup_n = int(node["poly_degree"])
            base = os.path.dirname(os.path.abspath(file_path))
            n = self.lookup_cells() * self.lookup_n
            for key, attr in (("depth_sigma_lookup", "depth_sigma_lookup"), ("depth_bias_lookup", "depth_bias_lookup")):
                with open(os.path.join(base, node[key])) as f:
                    vals = yaml.safe_load(f)
                setattr(self, attr, np.asarray(vals[:n], dtype=np.float32))
        elif "const_sigmasq_poly" in node:
            self.lookup_n = int(node["poly_degree"])
            self.const_sigmasq_poly = [float(v) for v in node["const_sigmasq_poly"][: self.lookup_n]]
            if "const_bias_poly" in node:
                self.const_bias_poly = [float(v) for v in node["const_bias_poly"][: self.lookup_n]]
            self.populate_lookup_tables()
        for key, conv in (("prior", float), ("occ_thresh", float), ("acc_thresh", float), ("max_iters", int)):
            if key in node:
                setattr(self, key, conv(node[key]))
        if "K" in node:
            self.K = np.asarray([float(v) for v in node["K"][:9]], dtype=np.float32).reshape(3, 3)
        if "gvdb_map_config" in node:
            self.gvdb_map_config = [int(v) for v in node["gvdb_map_config"][:5]]
        return self

    def set_from_python(self):
        """Make this the process-global parameter set (GVDBPyModules.cpp:25-29)."""
        global _global
        _global = self.copy()

    def copy(self):
        o = gvdb_params.__new__(gvdb_params)
        o.__dict__ = {k: (v.copy() if isinstance(v, (np.ndarray, list)) else v) for k, v in self.__dict__.items()}
        return o

    def to_c(self):
        """mrf_params for the C ABI; keeps the lookup arrays alive on the object."""
        p = _lib.MrfParams()
        p.K[:] = [float(v) for v in np.asarray(self.K, dtype=np.float32).reshape(9)]
        p.rows, p.cols = int(self.rows), int(self.cols)
        p.res, p.prior = float(self.res), float(self.prior)
        p.dims[:] = [float(v) for v in np.asarray(self.dims).reshape(3)]
        p.occ_thresh, p.acc_thresh = float(self.occ_thresh), float(self.acc_thresh)
        p.max_iters = int(self.max_iters)
        p.lookup_n = int(self.lookup_n)
        sp = list(self.const_sigmasq_poly) + [0.0] * (3 - len(self.const_sigmasq_poly))
        bp = list(self.const_bias_poly) + [0.0] * (3 - len(self.const_bias_poly))
        p.const_sigmasq_poly[:] = [float(v) for v in sp[:3]]
        p.const_bias_poly[:] = [float(v) for v in bp[:3]]
        if self.use_polys and self.depth_sigma_lookup is not None and self.depth_bias_lookup is not None:
            self._sig_keep = np.ascontiguousarray(self.depth_sigma_lookup, dtype=np.float32)
            self._bia_keep = np.ascontiguousarray(self.depth_bias_lookup, dtype=np.float32)
            p.depth_sigma_lookup = self._sig_keep.ctypes.data
            p.depth_bias_lookup = self._bia_keep.ctypes.data
        p.activation = int(self.activation)
        p.msg_dtype = int(self.msg_dtype)
        p.mask_848_border = int(bool(self.mask_848_border))
        p.brick_local_vis = int(getattr(self, "brick_local_vis", _lib.VIS_BRICK_LOCAL))
        return p

    def __repr__(self):
        return (f"gvdb_params(res={self.res}, prior={self.prior}, dims={list(np.asarray(self.dims))}, "
                f"rows={self.rows}, cols={self.cols}, max_iters={self.max_iters}, K={np.asarray(self.K).tolist()}, "
                f"sigma={self.const_sigmasq_poly}, lookup_n={self.lookup_n})")


_global = gvdb_params()


def global_params():
    return _global
