"""Batched persistent launch: many renders (different spins, inclinations,
distances and sizes) in one launch must give, render by render, the same
frame as a launch of that render alone.  Every ray's arithmetic is the same
in both, so FP64 and mixed frames are compared bit for bit, and the work
counters add up."""
import sys

import numpy as np
import pytest

from kerrtrace_amd import config as K

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _single(torch, cfg, cam):
    from kerrtrace_amd.renderer import HipRenderer
    with HipRenderer(cfg) as r:
        hdr, cls = r.trace(cam)
        torch.cuda.synchronize()
        return hdr.cpu().numpy(), cls.cpu().numpy(), r.stats()


def _renders(n, side, precision=K.KT_PRECISION_FP64, mixed_forms=False):
    out = []
    for k, (c, cam) in enumerate(K.batch_preset(limit=n, max_side=side)):
        # ragged sizes: make some renders non-square and not tile multiples
        if k % 3 == 1:
            c.width = max(9, c.width - 13)
        if k % 3 == 2:
            c.height = max(7, c.height // 2 + 5)
        c.precision = precision
        if mixed_forms and k % 2 == 1:
            # the reference's as-written right-hand side and secant root
            c.rhs_mode = K.KT_RHS_REFERENCE
            c.disk_root_mode = K.KT_DISK_ROOT_SECANT_REF
        out.append((c, cam))
    return out


@pytest.mark.parametrize("precision,mixed_forms", [(K.KT_PRECISION_FP64, False), (K.KT_PRECISION_MIXED, False),
                                                   (K.KT_PRECISION_FP64, True), (K.KT_PRECISION_MIXED, True)])
def test_batch_matches_single_renders(torch_gpu, precision, mixed_forms):
    """mixed_forms: every other render uses the as-written right-hand side
    and disk root; the batch kernel reads each render's form from its
    record, a single render's kernel is compiled for its form."""
    from kerrtrace_amd.renderer import HipBatch
    renders = _renders(8, 72, precision, mixed_forms)
    say(f"[batch] precision={precision} mixed_forms={mixed_forms} sizes={[(c.width, c.height) for c, _ in renders]}")
    with HipBatch(renders) as b:
        hdr, cls = b.trace()
        st = b.stats()
    attempts = 0
    for k, (c, cam) in enumerate(renders):
        s_hdr, s_cls, s_st = _single(torch_gpu, c, cam)
        attempts += s_st["attempts"]
        g_hdr, g_cls = hdr[k].cpu().numpy(), cls[k].cpu().numpy()
        assert np.array_equal(g_cls, s_cls), f"render {k}: hit classes differ"
        assert np.array_equal(g_hdr.view(np.uint32), s_hdr.view(np.uint32)), f"render {k}: HDR differs"
    assert st["rays"] == sum(c.width * c.height for c, _ in renders)
    assert st["attempts"] == attempts
    say(f"[batch] ok: {st['rays']} rays, {st['attempts']} attempts")


@pytest.mark.parametrize("opts", [(8, 8, 8, 2), (4, 16, 4, 1), (16, 1, 2, 0), (6, 32, 8, 1, 8, 1), (8, 16, 8, 1, 4, 2)])
def test_batch_launch_options_do_not_change_frames(torch_gpu, opts):
    from kerrtrace_amd.renderer import HipBatch
    renders = _renders(5, 64)
    say(f"[batch] opts={opts}")
    with HipBatch(renders) as b:
        ref, ref_cls = b.trace()
        ref = [t.cpu().numpy() for t in ref]
        b.set_launch_opts(*opts)
        hdr, cls = b.trace()
        for k in range(len(renders)):
            assert np.array_equal(hdr[k].cpu().numpy().view(np.uint32), ref[k].view(np.uint32)), f"render {k}"
            assert torch_gpu.equal(cls[k], ref_cls[k])


def test_batch_single_render_equals_renderer(torch_gpu):
    from kerrtrace_amd.renderer import HipBatch
    cfg, cam = K.scaled(*K.preset("hd1080_a0998"), 160, 90)
    with HipBatch([(cfg, cam)]) as b:
        hdr, cls = b.trace()
        g = hdr[0].cpu().numpy()
    s_hdr, s_cls, _ = _single(torch_gpu, cfg, cam)
    assert np.array_equal(g.view(np.uint32), s_hdr.view(np.uint32))
    assert np.array_equal(cls[0].cpu().numpy(), s_cls)


def test_batch_matches_oracle(torch_gpu):
    """Two small renders of the batch against the FP64 CPU oracle."""
    from kerrtrace_amd.renderer import HipBatch
    from oracle import oracle as O
    from tests.parity import compare
    renders = _renders(2, 48)
    with HipBatch(renders) as b:
        hdr, cls = b.trace()
        for k, (c, cam) in enumerate(renders):
            o_hdr, o_cls, _, _ = O.render(c, cam)
            m = compare(hdr[k].cpu().numpy(), cls[k].cpu().numpy(), o_hdr, o_cls)
            say(f"[batch] oracle render {k}: {m}")
            assert m["class_mismatch"] == 0
            assert m["exact_fp32_frac"] == 1.0


def test_autotune_keeps_frames_identical(torch_gpu):
    """The autotuner only picks launch options; the frame it leaves behind
    equals the default-options frame bit for bit."""
    from kerrtrace_amd.renderer import HipBatch, HipRenderer
    cfg, cam = K.scaled(*K.preset("hd1080_a0998"), 192, 108)
    with HipRenderer(cfg) as r:
        ref, _ = r.trace(cam)
        ref = ref.cpu().numpy()
        opts = r.autotune(cam)
        say(f"[autotune] single: {opts}")
        assert opts["waves_per_cu"] in (4, 6, 8) and opts["refill_threshold"] in (8, 16, 32)
        assert opts["isolate_level"] in (0, 1, 2)
        hdr, _ = r.trace(cam)
        assert np.array_equal(hdr.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    renders = _renders(4, 48)
    with HipBatch(renders) as b:
        ref, _ = b.trace()
        ref = [t.cpu().numpy() for t in ref]
        opts = b.autotune()
        say(f"[autotune] batch: {opts}")
        hdr, _ = b.trace()
        for k in range(len(renders)):
            assert np.array_equal(hdr[k].cpu().numpy().view(np.uint32), ref[k].view(np.uint32))


def test_batch_traces_on_two_streams(torch_gpu):
    """Consecutive batch traces on different streams are ordered by the
    batch itself (they share its device records and counters)."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipBatch
    renders = _renders(4, 48)
    say("[batch] two streams")
    with HipBatch(renders) as b:
        ref, _ = b.trace()
        ref = [t.cpu().numpy() for t in ref]
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        outs = []
        for k in range(4):
            hdr, _ = b.alloc(with_class=False)
            b.trace(hdr=hdr, with_class=False, stream=s1 if k % 2 == 0 else s2)
            outs.append(hdr)
        torch.cuda.synchronize()
        for hdr in outs:
            for k in range(len(renders)):
                assert np.array_equal(hdr[k].cpu().numpy().view(np.uint32), ref[k].view(np.uint32)), k
