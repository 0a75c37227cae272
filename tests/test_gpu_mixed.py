"""Mixed-precision tracer (FP32 stage bulk, FP64 radius solve / error norm /
controller / disk root / state) against the FP64 CPU oracle of the same
frame.  Bar: hit classes differ only on boundary pixels and on at most
0.1% of pixels; sub-classes (disk-contribution count, sky-grid line) differ
only on boundary pixels and on at most 0.2% of pixels; HDR relative error
<= 1e-4 on pixels of equal class away from class boundaries."""
import sys

import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O
from tests.parity import compare

pytestmark = pytest.mark.gpu

MIXED_MAX_MISMATCH_FRAC = 1e-3
MIXED_MAX_SUBCLASS_FRAC = 2e-3
MIXED_MAX_REL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu(torch, cfg, cam):
    from kerrtrace_amd.renderer import HipRenderer
    with HipRenderer(cfg) as r:
        hdr, cls = r.trace(cam)
        torch.cuda.synchronize()
        return hdr.cpu().numpy(), cls.cpu().numpy(), r.stats()


@pytest.mark.parametrize("name,size,tol", [
    ("cpu256_a09", (128, 128), 1e-8),
    ("hd1080_a0998", (384, 216), 1e-9),
    ("uhd2160_i60", (384, 216), 1e-7),
])
def test_mixed_matches_fp64_oracle(torch_gpu, name, size, tol):
    cfg, cam = K.preset(name)
    cfg, cam = K.scaled(cfg, cam, *size)
    cfg.tolerance = tol
    cfg.precision = K.KT_PRECISION_MIXED
    say(f"[mixed] {name} {size} tol={tol}")
    g_hdr, g_cls, st = _gpu(torch_gpu, cfg, cam)
    ref_cfg = cfg.copy()
    ref_cfg.precision = K.KT_PRECISION_FP64
    o_hdr, o_cls, _, ost = O.render(ref_cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[mixed] {name}: {m}; attempts gpu {st['attempts']} oracle {ost.attempts}")
    assert m["class_mismatch_non_boundary"] <= 2
    assert m["class_mismatch_frac"] <= MIXED_MAX_MISMATCH_FRAC
    assert m["subclass_mismatch_non_boundary"] <= 2
    assert m["subclass_mismatch"] <= MIXED_MAX_SUBCLASS_FRAC * m["pixels"]
    assert m["max_rel_non_boundary"] <= MIXED_MAX_REL
    # the split stage form keeps FP32 noise out of the error norm: the step
    # count stays close to the FP64 integration's
    assert st["attempts"] <= 1.15 * ost.attempts


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9, 0.998])
def test_mixed_spins_inclination_60(torch_gpu, spin):
    cfg, cam = K.preset("uhd2160_i60", spin=spin)
    cfg, cam = K.scaled(cfg, cam, 192, 108)
    say(f"[mixed] i=60 spin={spin}")
    g_hdr, g_cls, st = _gpu(torch_gpu, cfg, cam)
    ref = cfg.copy()
    ref.precision = K.KT_PRECISION_FP64
    o_hdr, o_cls, _, _ = O.render(ref, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[mixed] spin {spin}: {m}")
    assert m["class_mismatch_frac"] <= MIXED_MAX_MISMATCH_FRAC
    assert m["subclass_mismatch_non_boundary"] <= 2
    assert m["subclass_mismatch"] <= MIXED_MAX_SUBCLASS_FRAC * m["pixels"]
    assert m["max_rel_non_boundary"] <= MIXED_MAX_REL
