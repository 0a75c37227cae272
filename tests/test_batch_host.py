"""Host side of the batched launch: the committed 64-render draw matches its
generator, and kt_batch_create rejects bad input before touching a GPU."""
import ctypes as C
import importlib.util
import json
import os

import pytest

from kerrtrace_amd import config as K
from kerrtrace_amd._abi import KtCamera, KtConfig
from kerrtrace_amd._lib import last_error, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch64_fixture_matches_generator():
    path = os.path.join(ROOT, "tests", "golden", "make_batch64.py")
    spec = importlib.util.spec_from_file_location("make_batch64", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(K.BATCH64_PATH, encoding="utf-8") as fh:
        doc = json.load(fh)
    assert doc["renders"] == mod.draw()
    assert doc["total_rays"] == sum(r["side"] ** 2 for r in doc["renders"])
    sides = [r["side"] for r in doc["renders"]]
    assert len(sides) == 64 and 128 <= min(sides) and max(sides) <= 1024


def test_batch_preset_configs_validate():
    renders = K.batch_preset()
    assert len(renders) == 64
    for c, cam in renders:
        assert c.escape_radius == pytest.approx((cam.pos[0] ** 2 + cam.pos[1] ** 2 + cam.pos[2] ** 2) ** 0.5)
        assert lib().kt_validate_config(C.byref(c), None) == 0, last_error(None)


def _create(renders):
    n = len(renders)
    cfgs = (KtConfig * max(n, 1))(*[c for c, _ in renders])
    cams = (KtCamera * max(n, 1))(*[cam for _, cam in renders])
    h = C.c_void_p()
    return lib().kt_batch_create(cfgs, cams, n, C.byref(h)), h


def test_batch_create_rejects_bad_input():
    assert _create([])[0] != 0
    renders = K.batch_preset(limit=3, max_side=32)
    renders[1][0].precision = K.KT_PRECISION_MIXED
    rc, h = _create(renders)
    assert rc != 0 and not h.value
    assert "precision" in last_error(None)
    renders = K.batch_preset(limit=3, max_side=32)
    renders[2][0].width = 0
    rc, h = _create(renders)
    assert rc != 0 and not h.value
    assert "render 2" in last_error(None)
