"""The oracle's forward-mode gradient of the differentiable scene (observer
at theta = (spin, inclination [rad], r_obs, mdot)) against central finite
differences of the same integration with its step sequence frozen
(h = 1e-6 in every coordinate), in FP64 on the CPU.  The lookup tables are
piecewise linear, so a pixel whose temperature sits within the FD step of a
table node sees a kink: a small fraction of pixel-channels may differ; all
others agree to 1e-6, and the loss-level gradient agrees to 1e-5."""
import math

import numpy as np
import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O

FD_H = 1e-6
THETA = np.array([0.7, math.radians(70.0), 60.0, 1.0])


def _scene(size=40):
    cfg, _ = K.preset("grad512")
    cfg.width = cfg.height = size
    return cfg


@pytest.fixture(scope="module")
def dual_frame():
    cfg = _scene()
    return cfg, O.render_grad(cfg, THETA, K.GRAD_FOV_DEG)


def _central(cfg, k, h):
    tp = THETA.copy(); tp[k] += h
    tm = THETA.copy(); tm[k] -= h
    h0, hp, sp = O.render_frozen(cfg, THETA, tp, K.GRAD_FOV_DEG)
    _, hm, sm = O.render_frozen(cfg, THETA, tm, K.GRAD_FOV_DEG)
    return h0, (hp - hm) / (2 * h), sp & sm


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_dual_tangent_matches_frozen_central_difference(dual_frame, k):
    """Near-critical rays (photon-ring orbits) have derivatives of 1e3-1e4
    whose central difference is not converged at h = 1e-6; a pixel counts as
    resolved when the differences at h = 1e-6 and 1e-7 agree to 1e-5."""
    cfg, (hdr, tan, cls, _) = dual_frame
    h0, fd, same6 = _central(cfg, k, FD_H)
    _, fd7, same7 = _central(cfg, k, FD_H / 10)
    assert np.array_equal(h0, hdr)  # the dual render's values are the plain render
    d = tan[..., k]
    scale = np.abs(d).max()
    assert scale > 0
    floor = 1e-3 * np.median(np.abs(d[np.abs(d) > 0]))
    resolved = (same6 & same7)[..., None] & (np.abs(fd - fd7) <= 1e-5 * (np.abs(fd7) + floor))
    assert resolved.mean() > 0.97
    rel = np.abs(fd - d)[resolved] / (np.abs(d[resolved]) + floor)
    assert np.mean(rel > 1e-6) < 0.01
    # loss level over the resolved pixel-channels: L = sum w * hdr
    w = np.random.default_rng(k).uniform(-1, 1, hdr.shape) * resolved
    g_dual = float(np.sum(w * d))
    g_fd = float(np.sum(w * fd))
    assert g_fd == pytest.approx(g_dual, rel=1e-5, abs=1e-12 * scale * hdr.size)


def test_tangents_vanish_where_parameters_cannot_act(dual_frame):
    """mdot only rescales the disk temperature: rays that never composite a
    disk sample have zero d/dmdot; rays that end in the horizon without a
    disk sample have zero colour and zero tangents."""
    cfg, (hdr, tan, cls, _) = dual_frame
    no_disk = (cls & K.KT_FLAG_DISK) == 0
    assert no_disk.any() and (~no_disk).any()
    assert np.all(tan[no_disk][:, :, 3] == 0.0)
    dark = no_disk & ((cls & 3) == K.KT_END_HORIZON)
    assert np.all(hdr[dark] == 0.0) and np.all(tan[dark] == 0.0)
