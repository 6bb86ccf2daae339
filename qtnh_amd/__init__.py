"""qtnh_amd: MI355X (gfx950) native local contraction path of qtnh.

Layers:
  include/qtnh_hip.h   C ABI of the HIP kernels (libqtnh_hip.so)
  include/qtnh.hpp     C++ drop-in API of the reference (libqtnh.so)
  qtnh_amd.ops         Python front end of the C ABI over torch device memory
"""
from ._lib import (DIST_NORMAL, DIST_UNIFORM, PREC_FP64, PREC_SPLIT, X, QtnhError,  # noqa: F401
                   device_count, lib)

__all__ = ["lib", "device_count", "QtnhError", "X", "PREC_FP64", "PREC_SPLIT", "DIST_NORMAL", "DIST_UNIFORM"]
