"""kt_post_process_noalloc (bloom intermediate in the library's own device
array, no allocation) against kt_post_process (caller bloom buffer), through
the C ABI: bitwise equal images on two calls in a row, the second on a
smaller frame, and a frame above the array's size refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from kerrtrace_amd import config as K

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _frame(torch, cfg, seed):
    rng = np.random.default_rng(seed)
    h = np.ones((cfg.height, cfg.width, 4), dtype=np.float32)
    h[..., :3] = (rng.random((cfg.height, cfg.width, 3)) ** 3 * 4.0).astype(np.float32)
    return torch.from_numpy(h).cuda()


def _both(torch, r, hdr):
    """(image of kt_post_process, image of kt_post_process_noalloc)."""
    from kerrtrace_amd._lib import check, lib
    H, W = r.height, r.width
    bloom = torch.empty_like(hdr)
    ref = torch.empty((H, W), dtype=torch.int32, device="cuda")
    got = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    check(lib().kt_post_process(r._h, C.c_void_p(hdr.data_ptr()), C.c_void_p(bloom.data_ptr()),
                                C.c_void_p(ref.data_ptr()), C.c_void_p(s.cuda_stream)), r._h)
    check(lib().kt_post_process_noalloc(r._h, C.c_void_p(hdr.data_ptr()), C.c_void_p(got.data_ptr()),
                                        C.c_void_p(s.cuda_stream)), r._h)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_post_process_matches_post_process(torch_gpu):
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    # bloom sigma 12 px (R = 36) at 640x360, then sigma 8 px with a
    # bright-pass threshold at 200x120; the second call's smaller frame
    # reads only what the second horizontal pass wrote
    for (w, h), radius, thr, seed in [((640, 360), 36.0, 0.0, 1), ((200, 120), 24.0, 1.0, 2)]:
        cfg, _ = K.preset("cpu256_a09")
        cfg.width, cfg.height = w, h
        cfg.bloom_radius, cfg.bloom_threshold = radius, thr
        say(f"[noalloc] {w}x{h} R={radius} thr={thr}")
        with HipRenderer(cfg) as r:
            hdr = _frame(torch, cfg, seed)
            ref, got = _both(torch, r, hdr)
        assert np.array_equal(ref, got)
        assert len(np.unique(ref)) > 100


def test_noalloc_post_process_without_bloom_and_capacity(torch_gpu):
    torch = torch_gpu
    from kerrtrace_amd._abi import KT_ERR_CAPACITY
    from kerrtrace_amd._lib import lib
    from kerrtrace_amd.renderer import HipRenderer
    cfg, _ = K.preset("cpu256_a09")
    cfg.width, cfg.height = 96, 64
    cfg.bloom_enabled = 0
    say("[noalloc] bloom off 96x64")
    with HipRenderer(cfg) as r:
        ref, got = _both(torch, r, _frame(torch, cfg, 3))
    assert np.array_equal(ref, got)
    # 2048 x 1025 pixels exceed the 2,097,152-pixel array: refused before
    # any launch (the output stays untouched)
    cfg, _ = K.preset("cpu256_a09")
    cfg.width, cfg.height = 2048, 1025
    say("[noalloc] capacity check 2048x1025")
    with HipRenderer(cfg) as r:
        hdr = torch.zeros((cfg.height, cfg.width, 4), dtype=torch.float32, device="cuda")
        out = torch.full((cfg.height, cfg.width), -1, dtype=torch.int32, device="cuda")
        rc = lib().kt_post_process_noalloc(r._h, C.c_void_p(hdr.data_ptr()), C.c_void_p(out.data_ptr()),
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        assert rc == KT_ERR_CAPACITY
        assert bool((out == -1).all())
