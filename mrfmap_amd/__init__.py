"""mrfmap_amd — MI355X-native BP inference path of MRFMap.

Host-side mirror of the reference's Python module GVDBPyModules for the hot
path (src/GVDBPyModules.cpp:42-83) over libmrfmap_hip.so (C ABI in
include/mrfmap_hip.h; HIP kernels for gfx950 in mrfmap_amd/csrc/).
"""
from . import _lib  # noqa: F401
from ._lib import release_device_cache  # noqa: F401
from .params import gvdb_params, global_params  # noqa: F401
from .inference import GVDBImage, GVDBInference, GVDBMapLikelihoodEstimator  # noqa: F401

__all__ = ["gvdb_params", "global_params", "GVDBImage", "GVDBInference", "GVDBMapLikelihoodEstimator",
           "release_device_cache"]
