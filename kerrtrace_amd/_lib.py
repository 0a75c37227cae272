"""ctypes binding of libkerrtrace.so (include/kerrtrace.h).

The shared library is built in-tree by kerrtrace_amd/build.py.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

from ._abi import (KtCamera, KtConfig, KtDerived, KtFrameView, KtGradOpts,
                   KtGradStats, KtLaunchOpts, KtTraceStats)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkerrtrace.so")

# Every entry point include/kerrtrace.h declares, with its ctypes signature.
_i32, _u32, _f32, _f64, _vp = C.c_int32, C.c_uint32, C.c_float, C.c_double, C.c_void_p
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
SIGNATURES = {
    "kt_config_defaults": ([C.POINTER(KtConfig)], _i32),
    "kt_validate_config": ([C.POINTER(KtConfig), C.POINTER(KtDerived)], _i32),
    "kt_camera_from_yaw_pitch": ([_pd, _f64, _f64, _f64, C.POINTER(KtCamera)], _i32),
    "kt_camera_observer": ([_f64, _f64, _f64, C.POINTER(KtCamera)], _i32),
    "kt_last_error": ([_vp, C.c_char_p, C.c_size_t], _i32),
    "kt_host_luts": ([C.POINTER(KtConfig), _pd, _pd], _i32),
    "kt_create": ([C.POINTER(KtConfig), C.POINTER(_vp)], _i32),
    "kt_destroy": ([_vp], None),
    "kt_get_derived": ([_vp, C.POINTER(KtDerived)], _i32),
    "kt_submit_render": ([_vp, _pf, _pf, _pf, _pf, _f32, C.POINTER(_i32)], _i32),
    "kt_submit_render_f64": ([_vp, C.POINTER(KtCamera), C.POINTER(_i32)], _i32),
    "kt_ready_frame": ([_vp, C.POINTER(KtFrameView), C.POINTER(_i32)], _i32),
    "kt_finish_frame": ([_vp, _i32], _i32),
    "kt_trace": ([_vp, C.POINTER(KtCamera), _vp, _vp, _vp], _i32),
    "kt_trace_ex": ([_vp, C.POINTER(KtCamera), _vp, _vp, _vp, _vp], _i32),
    "kt_set_launch_opts": ([_vp, C.POINTER(KtLaunchOpts)], _i32),
    "kt_get_trace_stats": ([_vp, C.POINTER(KtTraceStats)], _i32),
    "kt_post_process": ([_vp, _vp, _vp, _vp, _vp], _i32),
    "kt_post_process_noalloc": ([_vp, _vp, _vp, _vp], _i32),
    "kt_get_luts": ([_vp, _pd, _pd], _i32),
    "kt_batch_create": ([C.POINTER(KtConfig), C.POINTER(KtCamera), _i32, C.POINTER(_vp)], _i32),
    "kt_batch_destroy": ([_vp], None),
    "kt_batch_set_launch_opts": ([_vp, C.POINTER(KtLaunchOpts)], _i32),
    "kt_batch_trace": ([_vp, C.POINTER(_vp), C.POINTER(_vp), _vp], _i32),
    "kt_batch_get_stats": ([_vp, C.POINTER(KtTraceStats)], _i32),
    "kt_autotune": ([_vp, C.POINTER(KtCamera), _vp, _vp, C.POINTER(KtLaunchOpts)], _i32),
    "kt_batch_autotune": ([_vp, C.POINTER(_vp), _vp, C.POINTER(KtLaunchOpts)], _i32),
    "kt_grad_create": ([C.POINTER(KtConfig), C.POINTER(KtGradOpts), C.POINTER(_vp)], _i32),
    "kt_grad_destroy": ([_vp], None),
    "kt_grad_forward": ([_vp, _pd, _vp, _vp, _vp], _i32),
    "kt_grad_backward": ([_vp, _vp, _vp, _vp, _vp], _i32),
    "kt_grad_get_stats": ([_vp, C.POINTER(KtGradStats)], _i32),
}

_lib = None


class KerrTraceError(RuntimeError):
    """A non-zero kt_status; ``code`` is the status value."""

    def __init__(self, code, message):
        super().__init__(f"kerrtrace error {code}: {message}")
        self.code = code
        self.message = message


def lib():
    """Load libkerrtrace.so once and attach the ctypes signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run kerrtrace_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error(handle=None):
    buf = C.create_string_buffer(512)
    lib().kt_last_error(handle, buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, handle=None):
    if rc != 0:
        raise KerrTraceError(rc, last_error(handle))
    return rc
