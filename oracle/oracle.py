"""ctypes front end of the FP64 CPU oracle (oracle/kt_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from kerrtrace_amd._abi import KtCamera, KtConfig, KtDerived

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libkt_oracle.so")
_lib = None


class KtoStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64), ("rhs_evals", C.c_uint64), ("attempts", C.c_uint64),
        ("accepted", C.c_uint64), ("crossings", C.c_uint64), ("steps", C.c_uint64),
        ("overflow", C.c_int32), ("reserved", C.c_int32), ("overflow_T", C.c_double),
        ("seconds", C.c_double), ("threads", C.c_int32), ("reserved2", C.c_int32),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        d, u64, u32, i32 = C.c_double, C.c_uint64, C.c_uint32, C.c_int
        pd = C.POINTER(C.c_double)
        L.kto_last_error.restype = C.c_char_p
        L.kto_pcg32.argtypes, L.kto_pcg32.restype = [u64], u32
        L.kto_rand01.argtypes, L.kto_rand01.restype = [u64], d
        L.kto_isco.argtypes, L.kto_isco.restype = [d, i32, d], d
        L.kto_nt_factor.argtypes, L.kto_nt_factor.restype = [d, d, d, d], d
        L.kto_pow_m02.argtypes, L.kto_pow_m02.restype = [d], d
        L.kto_turn_cos_sin.argtypes, L.kto_turn_cos_sin.restype = [u32, pd, pd], None
        L.kto_ks_r.argtypes, L.kto_ks_r.restype = [d, d, d, d], d
        L.kto_dense.argtypes, L.kto_dense.restype = [pd, d, d], d
        L.kto_dense_dt.argtypes, L.kto_dense_dt.restype = [pd, d, d], d
        L.kto_rhs_jvp.argtypes = [C.POINTER(KtConfig), pd, d, pd, d, d, pd, pd]
        L.kto_rhs_jvp.restype = C.c_int
        L.kto_step_jvp.argtypes = [C.POINTER(KtConfig), pd, d, d, pd, d, d, pd, pd]
        L.kto_step_jvp.restype = C.c_int
        L.kto_render_grad.argtypes = [C.POINTER(KtConfig), pd, d, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      pd, pd, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]
        L.kto_render_grad.restype = C.c_int
        L.kto_render_frozen.argtypes = [C.POINTER(KtConfig), pd, pd, d, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, pd, pd, C.POINTER(C.c_uint8)]
        L.kto_render_frozen.restype = C.c_int
        L.kto_disk_root.argtypes, L.kto_disk_root.restype = [pd, d], d
        L.kto_fbm.argtypes, L.kto_fbm.restype = [d, d, i32, u64], d
        L.kto_gradient_noise.argtypes, L.kto_gradient_noise.restype = [d, d, u64], d
        L.kto_derive.argtypes = [C.POINTER(KtConfig), C.POINTER(KtDerived), pd, pd]
        L.kto_get_derivs.argtypes = [C.POINTER(KtConfig), pd, d, pd]
        L.kto_init_ray.argtypes = [C.POINTER(KtConfig), pd, pd, pd, pd, pd]
        L.kto_render_rows.argtypes = [C.POINTER(KtConfig), C.POINTER(KtCamera), i32, i32, i32, i32,
                                      pd, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(KtoStats)]
        L.kto_post.argtypes = [C.POINTER(KtConfig), C.POINTER(C.c_float), C.POINTER(C.c_uint32), pd]
        L.kto_flop_counts.argtypes = [C.POINTER(KtConfig), pd, d, C.POINTER(C.c_uint64)]
        L.kto_dp_attempt.argtypes = [C.POINTER(KtConfig), pd, d, d, d, pd]
        L.kto_sky.argtypes = [C.POINTER(KtConfig), pd, d, pd, C.POINTER(C.c_int)]
        L.kto_step_tol.argtypes = [C.POINTER(KtConfig), pd, pd]
        L.kto_disk_redshift.argtypes = [C.POINTER(KtConfig), pd, d, d, pd, C.POINTER(C.c_int)]
        L.kto_fetch_disk.argtypes = [C.POINTER(KtConfig), d, pd]
        L.kto_ks_bl.argtypes = [C.POINTER(KtConfig), pd, pd]
        L.kto_pixel_dir.argtypes = [C.POINTER(KtConfig), C.POINTER(KtCamera), i32, i32, i32, pd]
        L.kto_fetch_blackbody.argtypes = [C.POINTER(KtConfig), d, pd, C.POINTER(C.c_int)]
        L.kto_trace_dir.argtypes = [C.POINTER(KtConfig), pd, pd, pd, C.POINTER(C.c_uint32), pd, i32,
                                    C.POINTER(C.c_int)]
        L.kto_trace_dir_x.argtypes = [C.POINTER(KtConfig), pd, pd, pd, pd, i32, C.POINTER(C.c_int), pd]
        for name in ("kto_derive", "kto_get_derivs", "kto_init_ray", "kto_render_rows", "kto_post",
                     "kto_flop_counts", "kto_dp_attempt", "kto_trace_dir", "kto_trace_dir_x", "kto_sky",
                     "kto_step_tol", "kto_disk_redshift", "kto_fetch_disk", "kto_fetch_blackbody",
                     "kto_ks_bl", "kto_pixel_dir"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().kto_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def pcg32(state):
    return lib().kto_pcg32(int(state) & 0xFFFFFFFFFFFFFFFF)


def rand01(state):
    return lib().kto_rand01(int(state) & 0xFFFFFFFFFFFFFFFF)


def isco(a_norm, prograde=True, mass=1.0):
    return lib().kto_isco(a_norm, 1 if prograde else 0, mass)


def nt_factor(r, a_norm, r_isco, inv_m):
    return lib().kto_nt_factor(r, a_norm, r_isco, inv_m)


def pow_m02(x):
    return lib().kto_pow_m02(x)


def turn_cos_sin(h):
    """(cos, sin) of the angle 2 pi h / 2^32, h a uint32."""
    c, s = C.c_double(), C.c_double()
    lib().kto_turn_cos_sin(int(h) & 0xFFFFFFFF, C.byref(c), C.byref(s))
    return c.value, s.value


def ks_r(x, y, z, a2):
    return lib().kto_ks_r(x, y, z, a2)


def dense(y0, y1, k1, k3, k4, k5, k6, k7, h, t, derivative=False):
    c = np.array([y0, y1, k1, k3, k4, k5, k6, k7], dtype=np.float64)
    f = lib().kto_dense_dt if derivative else lib().kto_dense
    return f(_dptr(c), h, t)


def disk_root(y0, y1, k1, k3, k4, k5, k6, k7, h):
    """True-root crossing parameter t in [0, 1] of the dense output's y."""
    c = np.array([y0, y1, k1, k3, k4, k5, k6, k7], dtype=np.float64)
    return lib().kto_disk_root(_dptr(c), h)


def rhs_jvp(cfg, state6, pt, dstate6, dpt, dspin):
    """Right-hand side and its derivative along (dstate, dpt, dspin)."""
    s = np.asarray(state6, dtype=np.float64); ds = np.asarray(dstate6, dtype=np.float64)
    out = np.zeros(6); tout = np.zeros(6)
    _check(lib().kto_rhs_jvp(C.byref(cfg), _dptr(s), pt, _dptr(ds), dpt, dspin, _dptr(out), _dptr(tout)))
    return out, tout


def step_jvp(cfg, state6, pt, h, dstate6, dpt, dspin):
    """End point of one DP5(4) step of size h and its derivative along
    (dstate, dpt, dspin) with h fixed."""
    s = np.asarray(state6, dtype=np.float64); ds = np.asarray(dstate6, dtype=np.float64)
    out = np.zeros(6); tout = np.zeros(6)
    _check(lib().kto_step_jvp(C.byref(cfg), _dptr(s), pt, h, _dptr(ds), dpt, dspin, _dptr(out), _dptr(tout)))
    return out, tout


def _block(cfg, block):
    return (0, 0, cfg.width, cfg.height) if block is None else tuple(int(v) for v in block)


def render_grad(cfg, theta, fov_deg, block=None, threads=0):
    """Differentiable scene (observer at theta = (a, i [rad], r_obs, mdot),
    escape radius r_obs) with forward-mode tangents: returns hdr [h, w, 3],
    tangents [h, w, 3, 4] (d hdr_c / d theta_k), cls [h, w], nsteps [h, w]."""
    x0, y0, w, h = _block(cfg, block)
    th = np.asarray(theta, dtype=np.float64)
    hdr = np.zeros((h, w, 3)); tan = np.zeros((h, w, 3, 4))
    cls = np.zeros((h, w), dtype=np.uint8); ns = np.zeros((h, w), dtype=np.uint32)
    rc = lib().kto_render_grad(C.byref(cfg), _dptr(th), fov_deg, x0, y0, w, h, threads if threads > 0 else host_threads(),
                               _dptr(hdr), _dptr(tan),
                               cls.ctypes.data_as(C.POINTER(C.c_uint8)), ns.ctypes.data_as(C.POINTER(C.c_uint32)))
    _check(rc)
    return hdr, tan, cls, ns


def render_frozen(cfg, theta0, theta1, fov_deg, block=None, threads=0):
    """Each pixel traced adaptively at theta0, then at theta1 replaying the
    same committed step sizes.  Returns hdr0, hdr1 [h, w, 3] and same [h, w]
    (True where both runs end with the same class and step count)."""
    x0, y0, w, h = _block(cfg, block)
    t0 = np.asarray(theta0, dtype=np.float64); t1 = np.asarray(theta1, dtype=np.float64)
    h0 = np.zeros((h, w, 3)); h1 = np.zeros((h, w, 3)); same = np.zeros((h, w), dtype=np.uint8)
    rc = lib().kto_render_frozen(C.byref(cfg), _dptr(t0), _dptr(t1), fov_deg, x0, y0, w, h,
                                 threads if threads > 0 else host_threads(), _dptr(h0),
                                 _dptr(h1), same.ctypes.data_as(C.POINTER(C.c_uint8)))
    _check(rc)
    return h0, h1, same.astype(bool)


def fbm(u, v, octaves, seed=0):
    return lib().kto_fbm(u, v, octaves, seed)


def gradient_noise(u, v, seed=0):
    return lib().kto_gradient_noise(u, v, seed)


def derive(cfg):
    """Derived parameters + (disk LUT, blackbody LUT) of a config."""
    out = KtDerived()
    disk = np.zeros(cfg.lut_size, dtype=np.float64)
    bb = np.zeros(cfg.lut_size * 3, dtype=np.float64)
    _check(lib().kto_derive(C.byref(cfg), C.byref(out), _dptr(disk), _dptr(bb)))
    return out, disk, bb.reshape(-1, 3)


def get_derivs(cfg, state6, pt):
    s = np.ascontiguousarray(state6, dtype=np.float64)
    out = np.zeros(6, dtype=np.float64)
    _check(lib().kto_get_derivs(C.byref(cfg), _dptr(s), pt, _dptr(out)))
    return out


def init_ray(cfg, cam_pos, direction):
    cam = np.ascontiguousarray(cam_pos, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    s = np.zeros(6, dtype=np.float64)
    pt, kt = C.c_double(), C.c_double()
    _check(lib().kto_init_ray(C.byref(cfg), _dptr(cam), _dptr(d), _dptr(s), C.byref(pt), C.byref(kt)))
    return s, pt.value, kt.value


MAX_THREADS = 16


def host_threads():
    """Worker threads for the oracle: the CPUs this process may run on,
    at most MAX_THREADS (a GPU host's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(MAX_THREADS, n))


def render_rows(cfg, cam, rows=None, threads=0):
    """Trace whole rows.  rows: None (all), or (begin, end, step);
    threads <= 0: host_threads().  Returns (hdr[nrows, W, 3] float64,
    cls[nrows, W] uint8, attempts[nrows, W] uint32, stats)."""
    if threads <= 0:
        threads = host_threads()
    if rows is None:
        rows = (0, cfg.height, 1)
    b, e, s = rows
    nrows = len(range(b, e, s))
    W = cfg.width
    hdr = np.zeros((nrows, W, 3), dtype=np.float64)
    cls = np.zeros((nrows, W), dtype=np.uint8)
    att = np.zeros((nrows, W), dtype=np.uint32)
    st = KtoStats()
    _check(lib().kto_render_rows(C.byref(cfg), C.byref(cam), b, e, s, threads, _dptr(hdr),
                                 cls.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 att.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(st)))
    return hdr, cls, att, st


def render(cfg, cam, threads=0):
    hdr, cls, att, st = render_rows(cfg, cam, None, threads)
    return hdr, cls, att, st


def post(cfg, hdr4):
    """LDR (uint32 0x00RRGGBB) and vertical-pass bloom of a float32 RGBA frame."""
    h = np.ascontiguousarray(hdr4, dtype=np.float32).reshape(cfg.height, cfg.width, 4)
    ldr = np.zeros((cfg.height, cfg.width), dtype=np.uint32)
    bloom = np.zeros((cfg.height, cfg.width, 3), dtype=np.float64)
    _check(lib().kto_post(C.byref(cfg), h.ctypes.data_as(C.POINTER(C.c_float)),
                          ldr.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(bloom)))
    return ldr, bloom


def flop_counts(cfg, state6, pt):
    """(FLOP per RHS evaluation, per attempt excluding RHS, per step commit)."""
    s = np.ascontiguousarray(state6, dtype=np.float64)
    out = (C.c_uint64 * 4)()
    _check(lib().kto_flop_counts(C.byref(cfg), _dptr(s), pt, out))
    return int(out[0]), int(out[1]), int(out[2])


def dp_attempt(cfg, state6, pt, h, tol):
    """One DP5(4) attempt: (k[7,6], next_state[6], error)."""
    s = np.ascontiguousarray(state6, dtype=np.float64)
    out = np.zeros(49, dtype=np.float64)
    _check(lib().kto_dp_attempt(C.byref(cfg), _dptr(s), pt, h, tol, _dptr(out)))
    return out[:42].reshape(7, 6), out[42:48].copy(), float(out[48])


def disk_redshift(cfg, x, z, px, py, pz, r, pt, kt):
    """(g, circular orbit exists) of a disk hit."""
    h = np.array([x, z, px, py, pz, r], dtype=np.float64)
    g = C.c_double()
    o = C.c_int()
    _check(lib().kto_disk_redshift(C.byref(cfg), _dptr(h), pt, kt, C.byref(g), C.byref(o)))
    return g.value, bool(o.value)


def pixel_dir(cfg, cam, x, y, sample=0):
    """The (not normalised) pinhole direction of a pixel's sample."""
    out = np.zeros(3, dtype=np.float64)
    _check(lib().kto_pixel_dir(C.byref(cfg), C.byref(cam), x, y, sample, _dptr(out)))
    return out


def ks_bl(cfg, x, y, z):
    """Boyer-Lindquist (r, sin theta, cos theta, phi) of a Kerr-Schild point."""
    p = np.array([x, y, z], dtype=np.float64)
    out = np.zeros(4, dtype=np.float64)
    _check(lib().kto_ks_bl(C.byref(cfg), _dptr(p), _dptr(out)))
    return tuple(float(v) for v in out)


def fetch_disk(cfg, u):
    """The disk factor lookup at u in [0, 1]."""
    out = C.c_double()
    _check(lib().kto_fetch_disk(C.byref(cfg), u, C.byref(out)))
    return out.value


def fetch_blackbody(cfg, T):
    """(rgb, overflow) of the blackbody lookup at temperature T."""
    rgb = np.zeros(3, dtype=np.float64)
    o = C.c_int()
    _check(lib().kto_fetch_blackbody(C.byref(cfg), T, _dptr(rgb), C.byref(o)))
    return rgb, bool(o.value)


def step_tol(cfg, state6):
    """Error tolerance of a step that starts at state6."""
    sv = np.ascontiguousarray(state6, dtype=np.float64)
    out = C.c_double()
    _check(lib().kto_step_tol(C.byref(cfg), _dptr(sv), C.byref(out)))
    return out.value


def sky(cfg, direction, transmittance=1.0):
    """Sky (intensity, on a grid line) along a unit direction."""
    dv = np.ascontiguousarray(direction, dtype=np.float64)
    out = C.c_double()
    g = C.c_int()
    _check(lib().kto_sky(C.byref(cfg), _dptr(dv), transmittance, C.byref(out), C.byref(g)))
    return out.value, bool(g.value)


def trace_dir(cfg, cam_pos, direction, max_log=0):
    """One ray: (rgb before exposure, class byte, stats dict, step log [n, 8])."""
    cam = np.ascontiguousarray(cam_pos, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    col = np.zeros(3, dtype=np.float64)
    st = (C.c_uint32 * 5)()
    log = np.zeros((max(max_log, 1), 8), dtype=np.float64)
    n = C.c_int(0)
    rc = lib().kto_trace_dir(C.byref(cfg), _dptr(cam), _dptr(d), _dptr(col), st,
                             _dptr(log) if max_log > 0 else None, max_log, C.byref(n))
    if rc < 0:
        _check(rc)
    stats = dict(zip(("attempts", "rhs", "accepted", "crossings", "steps"), list(st)))
    return col, rc, stats, log[:n.value].copy()


# Fields of a crossing record (kto_trace_dir_x), in order.
CROSS_FIELDS = ("t", "x", "z", "px", "py", "pz", "r", "in_disk", "orbit", "g", "u", "lut", "T_base", "T",
                "r_col", "g_col", "b_col", "alpha", "tr_before", "tr_after", "r_bl", "phi", "modulator",
                "composited")


def trace_dir_crossings(cfg, cam_pos, direction, max_cross=64):
    """One ray: (rgb before exposure, class byte, crossing records [n, 24]
    with columns CROSS_FIELDS, (pt, kt) of the ray)."""
    cam = np.ascontiguousarray(cam_pos, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    col = np.zeros(3, dtype=np.float64)
    rec = np.zeros((max(max_cross, 1), len(CROSS_FIELDS)), dtype=np.float64)
    pk = np.zeros(2, dtype=np.float64)
    n = C.c_int(0)
    rc = lib().kto_trace_dir_x(C.byref(cfg), _dptr(cam), _dptr(d), _dptr(col), _dptr(rec), max_cross,
                               C.byref(n), _dptr(pk))
    if rc < 0:
        _check(rc)
    return col, rc, rec[:n.value].copy(), (pk[0], pk[1])
