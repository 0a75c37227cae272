"""Python host over the C ABI, mirroring the reference's ``CudaRenderer``
(src/renderer/mod.rs:52-293): ``submit_render`` / ``ready_frame`` /
``finish_frame``, plus stream-ordered ``trace`` and ``post_process`` calls on
torch device tensors.  PyTorch only supplies device memory and streams.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._abi import (KtCamera, KtConfig, KtDerived, KtFrameView, KtLaunchOpts,
                   KtTraceStats)
from ._lib import check, lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("kerrtrace_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


@dataclass
class FrameView:
    """src/renderer/mod.rs:20-23: index + borrowed 0x00RRGGBB pixels."""
    index: int
    data: np.ndarray  # uint32 [H, W], valid until finish_frame(index)


class HipRenderer:
    """One renderer handle: two frame buffers with a stream each for the
    frame API, and stream-ordered traces on caller streams (traces issued on
    two streams run concurrently)."""

    def __init__(self, cfg: KtConfig):
        _torch()
        self.cfg = cfg.copy()
        h = C.c_void_p()
        check(lib().kt_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        d = KtDerived()
        check(lib().kt_get_derived(self._h, C.byref(d)), self._h)
        self.derived = d

    # ---- lifecycle
    def close(self):
        if getattr(self, "_h", None):
            lib().kt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def width(self):
        return self.cfg.width

    @property
    def height(self):
        return self.cfg.height

    # ---- stream-ordered device API
    def set_launch_opts(self, waves_per_cu=0, refill_threshold=0, queue_chunk=0, cost_order=0, long_ray_attempts=0,
                        isolate_level=0):
        """Persistent-tracer knobs (0 = default): resident waves per CU, idle
        lanes that trigger a refill, tile edge of the work order, tile order
        (0/1 hardest-first, 2 row-major), attempts after which a ray raises
        its wave's priority, and the priority level (1-3) from which such a
        wave takes no new rays (0: never)."""
        o = KtLaunchOpts(waves_per_cu, refill_threshold, queue_chunk, cost_order, long_ray_attempts, isolate_level)
        check(lib().kt_set_launch_opts(self._h, C.byref(o)), self._h)

    def autotune(self, cam, hdr=None, stream=None):
        """Pick the fastest launch options for this frame (kt_autotune);
        returns them as a dict.  hdr is overwritten."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        if hdr is None:
            with torch.cuda.stream(s):
                hdr = self.alloc_hdr()
        _check_tensor(hdr, (self.height, self.width, 4), torch.float32)
        o = KtLaunchOpts()
        check(lib().kt_autotune(self._h, C.byref(cam), C.c_void_p(hdr.data_ptr()), C.c_void_p(s.cuda_stream),
                                C.byref(o)), self._h)
        return {name: getattr(o, name) for name, _ in o._fields_}

    def alloc_hdr(self):
        torch = _torch()
        return torch.empty((self.height, self.width, 4), dtype=torch.float32, device="cuda")

    def trace(self, cam: KtCamera, hdr=None, cls=None, with_class=True, stream=None, attempts=None):
        """Trace one frame on `stream` (default: the current torch stream).
        Returns (hdr, cls): float32 [H, W, 4] RGBA (A = 1) and uint8 [H, W]
        hit classes; tensors it allocates belong to that stream.
        attempts: optional int32 [H, W] tensor receiving per-pixel attempts."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            if hdr is None:
                hdr = self.alloc_hdr()
            if cls is None and with_class:
                cls = torch.empty((self.height, self.width), dtype=torch.uint8, device="cuda")
        _check_tensor(hdr, (self.height, self.width, 4), torch.float32)
        if cls is not None:
            _check_tensor(cls, (self.height, self.width), torch.uint8)
        if attempts is not None:
            _check_tensor(attempts, (self.height, self.width), torch.int32)
            check(lib().kt_trace_ex(self._h, C.byref(cam), C.c_void_p(hdr.data_ptr()),
                                    C.c_void_p(cls.data_ptr() if cls is not None else 0),
                                    C.c_void_p(attempts.data_ptr()), C.c_void_p(s.cuda_stream)), self._h)
            return hdr, cls
        check(lib().kt_trace(self._h, C.byref(cam), C.c_void_p(hdr.data_ptr()),
                             C.c_void_p(cls.data_ptr() if cls is not None else 0),
                             C.c_void_p(s.cuda_stream)), self._h)
        return hdr, cls

    def stats(self):
        """Counters of the last trace (waits for it to finish)."""
        st = KtTraceStats()
        check(lib().kt_get_trace_stats(self._h, C.byref(st)), self._h)
        return {name: getattr(st, name) for name, _ in st._fields_}

    def post_process(self, hdr, ldr=None, bloom=None, stream=None):
        """Bloom + ACES + sRGB -> int32 [H, W] holding 0x00RRGGBB, on
        `stream` (default: the current torch stream).  bloom: float32
        [H, W, 4] scratch, allocated on that stream when not given."""
        torch = _torch()
        _check_tensor(hdr, (self.height, self.width, 4), torch.float32)
        s = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            if ldr is None:
                ldr = torch.empty((self.height, self.width), dtype=torch.int32, device="cuda")
            if bloom is None and self.derived.bloom_active:
                bloom = torch.empty_like(hdr)
        _check_tensor(ldr, (self.height, self.width), torch.int32)
        if bloom is not None:
            _check_tensor(bloom, (self.height, self.width, 4), torch.float32)
        check(lib().kt_post_process(self._h, C.c_void_p(hdr.data_ptr()),
                                    C.c_void_p(bloom.data_ptr() if bloom is not None else 0),
                                    C.c_void_p(ldr.data_ptr()), C.c_void_p(s.cuda_stream)), self._h)
        return ldr

    def luts(self):
        n = self.cfg.lut_size
        disk = np.zeros(n, dtype=np.float64)
        bb = np.zeros(n * 3, dtype=np.float64)
        check(lib().kt_get_luts(self._h, disk.ctypes.data_as(C.POINTER(C.c_double)),
                                bb.ctypes.data_as(C.POINTER(C.c_double))), self._h)
        return disk, bb.reshape(-1, 3)

    # ---- frame API (src/renderer/mod.rs:119-260)
    def submit_render(self, cam_pos, fwd, rgt, up, fov_scale):
        f3 = C.c_float * 3
        ok = C.c_int32(0)
        check(lib().kt_submit_render(self._h, f3(*cam_pos), f3(*fwd), f3(*rgt), f3(*up),
                                     C.c_float(fov_scale), C.byref(ok)), self._h)
        return bool(ok.value)

    def submit_render_camera(self, cam: KtCamera):
        ok = C.c_int32(0)
        check(lib().kt_submit_render_f64(self._h, C.byref(cam), C.byref(ok)), self._h)
        return bool(ok.value)

    def ready_frame(self):
        v = KtFrameView()
        has = C.c_int32(0)
        check(lib().kt_ready_frame(self._h, C.byref(v), C.byref(has)), self._h)
        if not has.value:
            return None
        n = v.width * v.height
        arr = np.ctypeslib.as_array(v.data, shape=(n,)).reshape(v.height, v.width)
        return FrameView(index=v.index, data=arr)

    def finish_frame(self, index):
        check(lib().kt_finish_frame(self._h, int(index)), self._h)


class HipBatch:
    """Many renders traced by one persistent launch (kt_batch_*): each keeps
    its own config and camera; their tiles share one hardest-first work
    queue.  All renders use the same precision mode."""

    def __init__(self, renders):
        _torch()
        self.cfgs = [c.copy() for c, _ in renders]
        self.cams = [cam.copy() for _, cam in renders]
        n = len(self.cfgs)
        cfg_arr = (KtConfig * n)(*self.cfgs)
        cam_arr = (KtCamera * n)(*self.cams)
        h = C.c_void_p()
        check(lib().kt_batch_create(cfg_arr, cam_arr, n, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().kt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.cfgs)

    @property
    def rays(self):
        return sum(c.width * c.height * c.spp for c in self.cfgs)

    def set_launch_opts(self, waves_per_cu=0, refill_threshold=0, queue_chunk=0, cost_order=0, long_ray_attempts=0,
                        isolate_level=0):
        o = KtLaunchOpts(waves_per_cu, refill_threshold, queue_chunk, cost_order, long_ray_attempts, isolate_level)
        check(lib().kt_batch_set_launch_opts(self._h, C.byref(o)))

    def autotune(self, hdr=None, stream=None):
        """Pick the fastest launch options for the batch (kt_batch_autotune);
        returns them as a dict.  The hdr tensors are overwritten."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        if hdr is None:
            with torch.cuda.stream(s):
                hdr, _ = self.alloc(with_class=False)
        n = len(self.cfgs)
        if len(hdr) != n:
            raise ValueError(f"expected {n} output tensors")
        for i, c in enumerate(self.cfgs):
            _check_tensor(hdr[i], (c.height, c.width, 4), torch.float32)
        hp = (C.c_void_p * n)(*[t.data_ptr() for t in hdr])
        o = KtLaunchOpts()
        check(lib().kt_batch_autotune(self._h, hp, C.c_void_p(s.cuda_stream), C.byref(o)))
        return {name: getattr(o, name) for name, _ in o._fields_}

    def alloc(self, with_class=True):
        torch = _torch()
        hdr = [torch.empty((c.height, c.width, 4), dtype=torch.float32, device="cuda") for c in self.cfgs]
        cls = [torch.empty((c.height, c.width), dtype=torch.uint8, device="cuda") for c in self.cfgs] \
            if with_class else None
        return hdr, cls

    def trace(self, hdr=None, cls=None, with_class=True, stream=None):
        """Trace every render on the current torch stream; returns lists of
        float32 [H, W, 4] and (optionally) uint8 [H, W] tensors."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        if hdr is None:
            with torch.cuda.stream(s):
                hdr, c2 = self.alloc(with_class and cls is None)
            cls = cls if cls is not None else c2
        n = len(self.cfgs)
        if len(hdr) != n or (cls is not None and len(cls) != n):
            raise ValueError(f"expected {n} output tensors per list")
        for i, c in enumerate(self.cfgs):
            _check_tensor(hdr[i], (c.height, c.width, 4), torch.float32)
            if cls is not None:
                _check_tensor(cls[i], (c.height, c.width), torch.uint8)
        hp = (C.c_void_p * n)(*[t.data_ptr() for t in hdr])
        cp = (C.c_void_p * n)(*[t.data_ptr() for t in cls]) if cls is not None else None
        check(lib().kt_batch_trace(self._h, hp, cp, C.c_void_p(s.cuda_stream)))
        return hdr, cls

    def stats(self):
        st = KtTraceStats()
        check(lib().kt_batch_get_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}


def _check_tensor(t, shape, dtype):
    """A caller's buffer goes to the kernels as a raw pointer: reject any
    that is not exactly the contiguous device tensor they index."""
    if (not hasattr(t, "shape") or tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_cuda
            or not t.is_contiguous()):
        got = f"{tuple(t.shape)} {t.dtype}" if hasattr(t, "shape") else type(t).__name__
        raise ValueError(f"expected a contiguous cuda tensor {shape} {dtype}, got {got}")


def unpack_ldr(ldr):
    """0x00RRGGBB (int32/uint32 array) -> uint8 [H, W, 3]."""
    a = np.asarray(ldr).astype(np.uint32)
    return np.stack([(a >> 16) & 255, (a >> 8) & 255, a & 255], axis=-1).astype(np.uint8)
