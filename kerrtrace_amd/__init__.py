"""kerrtrace_amd: MI355X-native (gfx950 HIP) Kerr black-hole geodesic renderer.

The renderer lives in libkerrtrace.so (HIP kernels + C++ runtime behind the
C ABI in include/kerrtrace.h).  This package holds the Python side of that
boundary: the ctypes binding, the config.yaml loader and presets, and the
frame/trace API mirroring the reference's CudaRenderer.
"""
from ._abi import (KT_DISK_ROOT_SECANT_REF, KT_DISK_ROOT_TRUE, KT_END_ESCAPE,
                   KT_END_HORIZON, KT_END_OTHER, KT_FLAG_DISK,
                   KT_FLAG_LUT_OVERFLOW, KT_FLAG_SKY_GRID, KT_PRECISION_FP64,
                   KT_PRECISION_MIXED, KtCamera, KtConfig)
from .config import (camera_from_yaw_pitch, default_config, load_config,
                     observer_camera, preset)

__all__ = [
    "KtConfig", "KtCamera", "default_config", "load_config", "preset",
    "observer_camera", "camera_from_yaw_pitch", "KT_PRECISION_FP64",
    "KT_PRECISION_MIXED", "KT_DISK_ROOT_SECANT_REF", "KT_DISK_ROOT_TRUE",
    "KT_END_OTHER", "KT_END_HORIZON", "KT_END_ESCAPE", "KT_FLAG_DISK",
    "KT_FLAG_SKY_GRID", "KT_FLAG_LUT_OVERFLOW",
]


def HipRenderer(*args, **kwargs):
    """Create a renderer (loads libkerrtrace.so; needs a GPU)."""
    from .renderer import HipRenderer as _R
    return _R(*args, **kwargs)
