"""The adjoint pass's building blocks (kerrtrace_amd/csrc/kt_adjoint.hpp),
built for the host from the same source, against the FP64 oracle on the
CPU: right-hand-side and step VJPs against the oracle's forward-mode
dual-number JVPs (w . J v computed both ways), the crossing's reverse
against central differences, the ray start and the emission's dual-number
tangents against central differences of their values."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "kt_adjoint_check.cpp")
HDRS = [os.path.join(HERE, "..", "kerrtrace_amd", "csrc", f) for f in ("kt_adjoint.hpp", "kt_math.hpp",
                                                                          "kt_internal.hpp")]
OUT = os.path.join(HERE, "native", "_build", "libkt_adjoint_check.so")

_pd = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kta():
    stale = not os.path.exists(OUT) or any(os.path.getmtime(f) > os.path.getmtime(OUT) for f in [SRC] + HDRS)
    if stale:
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", SRC, "-o", OUT]
        subprocess.run(cmd, check=True, capture_output=True)
    L = C.CDLL(OUT)
    d, i = C.c_double, C.c_int
    L.kta_rhs_vjp.argtypes = [_pd, d, d, d, _pd, _pd, _pd, _pd]
    L.kta_step.argtypes = [_pd, d, d, d, d, _pd]
    L.kta_rhs_vjp32.argtypes = [_pd, d, d, d, _pd, _pd, _pd, _pd]
    L.kta_step_reverse32.argtypes = [_pd, d, d, d, d, _pd, _pd, _pd, _pd]
    L.kta_step_reverse.argtypes = [_pd, d, d, d, d, _pd, _pd, _pd, _pd]
    L.kta_crossing.argtypes = [_pd, _pd, _pd, d, i, _pd, _pd, _pd, _pd, _pd]
    L.kta_start.argtypes = [i, i, i, i, d, d, d, d, d, i, _pd, _pd]
    L.kta_shade.argtypes = [_pd, C.POINTER(C.c_int), C.c_uint64, _pd, _pd, _pd, _pd, d, d, i, _pd, _pd]
    L.kta_stages32.argtypes = [_pd, d, d, d, d, _pd, _pd]
    L.kta_pow_m02.argtypes = [d]
    L.kta_pow_m02.restype = d
    return L


@pytest.mark.parametrize("spin", [0.0, 0.6, 0.998])
def test_fp32_stage_inputs_match_fp64(kta, spin):
    """The FP32 reversal's stage inputs, formed again in FP32 (dp_stages32:
    FP32 right-hand side including the radius solve, FP32 stage sums), lie
    within FP32 rounding of the FP64 stage inputs of the same step: the
    linearisation points of the FP32 VJPs (measured 2.4e-6 of the state's
    scale per component: FP32 rounding amplified by the tableau's weights,
    up to 11.6 in magnitude).  A wrong tableau entry or a dropped term in
    the FP32 right-hand side misses the bar (1e-5) by orders of magnitude."""
    rng = np.random.default_rng(11)
    y64, y32 = np.zeros(30), np.zeros(30)
    worst = 0.0
    for _ in range(200):
        s, pt = _state(rng, spin)
        h = float(rng.uniform(0.05, 2.0))
        kta.kta_stages32(_p(s), h, pt, spin, 1.0, _p(y64), _p(y32))
        Y64, Y32 = y64.reshape(5, 6), y32.reshape(5, 6)
        scale = np.maximum(np.abs(Y64), 1.0)
        worst = max(worst, float(np.max(np.abs(Y32 - Y64) / scale)))
    assert worst <= 1e-5, worst


def test_device_pow_m02_matches_oracle(kta):
    """The device's branch-free controller factor (special values selected
    after the main path, q = floor(e/5) from an unsigned division) returns
    the oracle's bits for every class of input: NaN, +-0, +-inf, negatives,
    subnormals, the extremes and 20000 log-uniform positives.  A wrong
    special-value select, or a q off by one for negative exponents (e.g. a
    truncating division), fails here."""
    rng = np.random.default_rng(5)
    xs = [math.nan, 0.0, -0.0, math.inf, -math.inf, -1.0, -1e-300, 5e-324, 1e-310, 2.2250738585072014e-308,
          1.7976931348623157e308, 1.0, 2.0, 0.5, 1e-9, 3e-5]
    xs += [2.0 ** k for k in range(-1074, 1024, 7)]
    xs += list(np.exp(rng.uniform(-700.0, 700.0, 20000)))
    for x in xs:
        got, want = kta.kta_pow_m02(float(x)), O.pow_m02(float(x))
        if math.isnan(want):
            assert math.isnan(got), x
        else:
            assert np.float64(got).tobytes() == np.float64(want).tobytes(), (x, got, want)


def _p(a):
    return a.ctypes.data_as(_pd)


def _cfg(spin, mass=1.0):
    cfg, _ = K.preset("grad512", spin=spin)
    cfg.mass = mass
    return cfg


def _state(rng, a):
    """A photon-like state at 3..30 M off the equator, not on the spin axis."""
    r = rng.uniform(3.0, 30.0)
    th = rng.uniform(0.3, 2.8)
    ph = rng.uniform(0, 2 * math.pi)
    x = np.array([r * math.sin(th) * math.cos(ph), r * math.cos(th), r * math.sin(th) * math.sin(ph)])
    p = rng.normal(size=3)
    p /= np.linalg.norm(p)
    return np.concatenate([x, p]), -1.0 - 0.1 * rng.uniform()


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9, 0.998, -0.3])
def test_rhs_vjp_matches_oracle_jvp(kta, spin):
    rng = np.random.default_rng(int(1000 * (spin + 1)))
    cfg = _cfg(spin)
    for _ in range(20):
        s, pt = _state(rng, spin)
        w = rng.normal(size=6)
        lam = np.zeros(6)
        lpt, la = C.c_double(), C.c_double()
        kta.kta_rhs_vjp(_p(s), pt, spin, 1.0, _p(w), _p(lam), C.byref(lpt), C.byref(la))
        for _ in range(3):
            ds, dpt, da = rng.normal(size=6), rng.normal(), rng.normal()
            _, tout = O.rhs_jvp(cfg, s, pt, ds, dpt, da)
            lhs = float(w @ tout)
            terms = np.concatenate([lam * ds, [lpt.value * dpt, la.value * da]])
            assert float(terms.sum()) == pytest.approx(lhs, rel=1e-10, abs=1e-12 * np.abs(terms).sum())


@pytest.mark.parametrize("spin", [0.0, 0.7, 0.998])
def test_step_matches_oracle_and_reverse_matches_jvp(kta, spin):
    rng = np.random.default_rng(int(77 + 100 * spin))
    cfg = _cfg(spin)
    for _ in range(10):
        s, pt = _state(rng, spin)
        h = float(rng.uniform(0.05, 1.0))
        ns = np.zeros(6)
        kta.kta_step(_p(s), h, pt, spin, 1.0, _p(ns))
        lam_ns = rng.normal(size=6)
        lam_s = np.zeros(6)
        lpt, la = C.c_double(), C.c_double()
        kta.kta_step_reverse(_p(s), h, pt, spin, 1.0, _p(lam_ns), _p(lam_s), C.byref(lpt), C.byref(la))
        for _ in range(3):
            ds, dpt, da = rng.normal(size=6), rng.normal(), rng.normal()
            out, tout = O.step_jvp(cfg, s, pt, h, ds, dpt, da)
            # the stepping arithmetic is the tracer's (bit for bit the oracle's)
            assert np.array_equal(out, ns)
            lhs = float(lam_ns @ tout)
            terms = np.concatenate([lam_s * ds, [lpt.value * dpt, la.value * da]])
            assert float(terms.sum()) == pytest.approx(lhs, rel=1e-10, abs=1e-12 * np.abs(terms).sum())


@pytest.mark.parametrize("spin", [0.0, 0.7, 0.998])
def test_fp32_vjp_and_step_reverse_match_oracle_jvp(kta, spin):
    """The FP32 reversal's building blocks against the oracle's FP64 JVP:
    <w, J v> = <J^T w, v> to FP32 accuracy (relative 1e-5 of the terms'
    magnitudes: a few ulp of float over a dot product of 8 terms).  A
    wrong sign or index in the FP32 dual pass fails at O(1)."""
    rng = np.random.default_rng(int(31 + 100 * spin))
    cfg = _cfg(spin)
    for _ in range(10):
        s, pt = _state(rng, spin)
        w = rng.normal(size=6)
        lam = np.zeros(6)
        lpt, la = C.c_double(), C.c_double()
        kta.kta_rhs_vjp32(_p(s), pt, spin, 1.0, _p(w), _p(lam), C.byref(lpt), C.byref(la))
        h = float(rng.uniform(0.05, 1.0))
        lam_ns = rng.normal(size=6)
        lam_s = np.zeros(6)
        spt, sa = C.c_double(), C.c_double()
        kta.kta_step_reverse32(_p(s), h, pt, spin, 1.0, _p(lam_ns), _p(lam_s), C.byref(spt), C.byref(sa))
        for _ in range(3):
            ds, dpt, da = rng.normal(size=6), rng.normal(), rng.normal()
            _, tout = O.rhs_jvp(cfg, s, pt, ds, dpt, da)
            terms = np.concatenate([lam * ds, [lpt.value * dpt, la.value * da]])
            assert float(terms.sum()) == pytest.approx(float(w @ tout), rel=1e-5, abs=1e-5 * np.abs(terms).sum())
            _, tstep = O.step_jvp(cfg, s, pt, h, ds, dpt, da)
            terms = np.concatenate([lam_s * ds, [spt.value * dpt, sa.value * da]])
            assert float(terms.sum()) == pytest.approx(float(lam_ns @ tstep), rel=1e-5,
                                                       abs=1e-5 * np.abs(terms).sum())


@pytest.mark.parametrize("root_mode", [K.KT_DISK_ROOT_TRUE, K.KT_DISK_ROOT_SECANT_REF])
def test_crossing_reverse_matches_central_differences(kta, root_mode):
    rng = np.random.default_rng(3 + root_mode)
    for _ in range(5):
        s = np.array([8.0, 0.3, 1.0, -0.2, -0.9, 0.1]) + 0.05 * rng.normal(size=6)
        ns = s + np.array([-0.2, -0.6, 0.05, -0.01, -0.02, 0.0]) + 0.01 * rng.normal(size=6)
        k = rng.normal(size=42) * 0.3
        k[1] = k[37] = -0.6  # y rates of k1 and k7 consistent with a crossing
        h = 1.0
        lam_hit = rng.normal(size=5)
        hit = np.zeros(6)
        ls, lns, lk = np.zeros(6), np.zeros(6), np.zeros(36)
        kta.kta_crossing(_p(s), _p(ns), _p(k), h, root_mode, _p(lam_hit), _p(hit), _p(ls), _p(lns), _p(lk))
        assert 0.0 < hit[0] < 1.0

        def f(s_, ns_, k_):
            out = np.zeros(6)
            a, b, c = np.zeros(6), np.zeros(6), np.zeros(36)
            kta.kta_crossing(_p(s_), _p(ns_), _p(k_), h, root_mode, _p(lam_hit), _p(out), _p(a), _p(b), _p(c))
            return float(lam_hit @ out[1:])

        eps = 1e-6
        for j in range(6):
            e = np.zeros(6); e[j] = eps
            fd = (f(s + e, ns, k) - f(s - e, ns, k)) / (2 * eps)
            assert ls[j] == pytest.approx(fd, rel=1e-6, abs=1e-7)
            fd = (f(s, ns + e, k) - f(s, ns - e, k)) / (2 * eps)
            assert lns[j] == pytest.approx(fd, rel=1e-6, abs=1e-7)
        for slot, idx in enumerate([0, 2, 3, 4, 5, 6]):  # k1, k3..k7
            for j in range(6):
                e = np.zeros(42); e[6 * idx + j] = eps
                fd = (f(s, ns, k + e) - f(s, ns, k - e)) / (2 * eps)
                assert lk[6 * slot + j] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def _observer_dir(W, H, px, py, incl, fov_scale):
    base = (py * W + px) << 32
    rx, ry = O.rand01(base), O.rand01(base ^ 0xda3e39cb94b95bdb)
    u = ((px + rx) * (2.0 / W) - 1.0) * (W / H) * fov_scale
    v = (1.0 - (py + ry) * (2.0 / H)) * fov_scale
    si, ci = math.sin(incl), math.cos(incl)
    return np.array([v * -ci - si, v * si - ci, -u])


def test_ray_start_values_and_tangents(kta):
    W, H = 40, 30
    fov_scale = math.tan(math.radians(K.GRAD_FOV_DEG) / 2)
    th = np.array(K.GRAD_TARGET)
    cfg = _cfg(th[0])
    for (px, py) in [(0, 0), (13, 21), (39, 29), (20, 15)]:
        vals, tans = np.zeros(8), np.zeros(8)
        kta.kta_start(W, H, px, py, th[0], 1.0, th[1], th[2], fov_scale, 0, _p(vals), _p(tans))
        si, ci = math.sin(th[1]), math.cos(th[1])
        s0, pt, kt = O.init_ray(cfg, [th[2] * si, th[2] * ci, 0.0], _observer_dir(W, H, px, py, th[1], fov_scale))
        np.testing.assert_allclose(vals, np.concatenate([s0, [pt, kt]]), rtol=1e-13, atol=1e-13)
        for d in range(3):
            kta.kta_start(W, H, px, py, th[0], 1.0, th[1], th[2], fov_scale, d, _p(vals), _p(tans))
            eps = 1e-6
            tp, tm = th.copy(), th.copy()
            tp[d] += eps
            tm[d] -= eps
            vp, vm, junk = np.zeros(8), np.zeros(8), np.zeros(8)
            kta.kta_start(W, H, px, py, tp[0], 1.0, tp[1], tp[2], fov_scale, d, _p(vp), _p(junk))
            kta.kta_start(W, H, px, py, tm[0], 1.0, tm[1], tm[2], fov_scale, d, _p(vm), _p(junk))
            np.testing.assert_allclose(tans, (vp - vm) / (2 * eps), rtol=1e-6, atol=1e-8)


def _shade_setup(spin, mdot=1.0):
    cfg = _cfg(spin)
    cfg.mdot_scale = mdot
    d, disk, bb = O.derive(cfg)
    return cfg, d, disk, bb


def test_shade_tangents_match_central_differences(kta):
    """The emission's dual-number pass against central differences of its
    own values; in spin and mdot the LUT, ISCO and temperature scale are
    re-derived by the oracle at each shifted parameter."""
    th = K.GRAD_TARGET
    spin, mdot = th[0], th[3]
    cfg, d, disk, bb = _shade_setup(spin, mdot)
    eps = 1e-6
    cp, dp_, diskp, _ = _shade_setup(spin + eps, mdot)
    cm, dm_, diskm, _ = _shade_setup(spin - eps, mdot)
    dlut = (diskp - diskm) / (2 * eps)
    d_rin = (dp_.disk_inner - dm_.disk_inner) / (2 * eps)
    dTs_da = (dp_.temperature_scale - dm_.temperature_scale) / (2 * eps)
    _, dmp, _, _ = _shade_setup(spin, mdot + eps)
    _, dmm, _, _ = _shade_setup(spin, mdot - eps)
    dTs_dm = (dmp.temperature_scale - dmm.temperature_scale) / (2 * eps)
    ip = (C.c_int * 2)(cfg.lut_size, cfg.noise_detail)
    seed_mix = (cfg.noise_seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF

    def sp_of(a, dd):
        return np.array([a, cfg.mass, dd.disk_inner, dd.disk_outer, dd.temperature_scale, cfg.lut_max_temp,
                         cfg.wavelength_step, cfg.noise_scale, cfg.noise_strength, cfg.noise_winding, d_rin, dTs_da,
                         dTs_dm])

    def run(sp, lut, hit, pt, kt, direction):
        vals, tans = np.zeros(5), np.zeros(4)
        kta.kta_shade(_p(sp), ip, seed_mix, _p(lut), _p(dlut), _p(np.ascontiguousarray(bb.ravel())), _p(hit), pt,
                      kt, direction, _p(vals), _p(tans))
        return vals, tans

    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(30):
        r = rng.uniform(d.disk_inner + 0.5, d.disk_outer - 0.5)
        ph = rng.uniform(0, 2 * math.pi)
        hit = np.array([r * math.cos(ph), r * math.sin(ph), rng.normal(), -0.5 + 0.1 * rng.normal(), rng.normal()])
        pt, kt = -1.0 - 0.05 * rng.uniform(), -1.0 - 0.05 * rng.uniform()
        sp = sp_of(spin, d)
        vals, _ = run(sp, disk, hit, pt, kt, 0)
        if vals[4] != 1.0:
            continue
        for direction in range(9):
            _, tans = run(sp, disk, hit, pt, kt, direction)
            if direction < 7:
                e = np.zeros(7); e[direction] = eps
                hp, hm = hit + e[:5], hit - e[:5]
                vp, _ = run(sp, disk, hp, pt + e[5], kt + e[6], direction)
                vm, _ = run(sp, disk, hm, pt - e[5], kt - e[6], direction)
            elif direction == 7:
                vp, _ = run(sp_of(spin + eps, dp_), diskp, hit, pt, kt, direction)
                vm, _ = run(sp_of(spin - eps, dm_), diskm, hit, pt, kt, direction)
            else:
                spp, spm = sp.copy(), sp.copy()
                spp[4], spm[4] = dmp.temperature_scale, dmm.temperature_scale
                vp, _ = run(spp, disk, hit, pt, kt, direction)
                vm, _ = run(spm, disk, hit, pt, kt, direction)
            if vp[4] != 1.0 or vm[4] != 1.0:
                continue
            fd = (vp[:4] - vm[:4]) / (2 * eps)
            scale = np.abs(fd).max() + np.abs(tans).max() + 1e-12
            # LUT cells are linear pieces: a step across a node is a kink
            assert np.abs(tans - fd).max() <= 1e-4 * scale, (direction, tans, fd)
        checked += 1
    assert checked >= 10
