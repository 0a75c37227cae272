"""The C ABI: libkerrtrace.so loads, exports every entry point that
include/kerrtrace.h declares, and its host-only functions (config
defaults and validation, host lookup tables, cameras) agree with the
reference's host logic and the oracle.  No GPU calls here."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from kerrtrace_amd import _abi, config as K
from kerrtrace_amd._lib import SIGNATURES, KerrTraceError, check, lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kerrtrace.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(kt_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    declared = _declared_functions()
    assert declared, "no declarations parsed"
    L = lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(SIGNATURES), declared ^ set(SIGNATURES)


def _c_struct_sizes():
    """sizeof/offsetof of the header's structs, from a tiny C program."""
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "kerrtrace.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(kt_config), sizeof(kt_camera), sizeof(kt_derived),
         sizeof(kt_trace_stats), sizeof(kt_launch_opts), sizeof(kt_frame_view));
  printf("%zu %zu %zu %zu\n", offsetof(kt_config, noise_seed), offsetof(kt_config, bloom_threshold),
         offsetof(kt_config, precision), offsetof(kt_config, rhs_mode));
  printf("%zu %zu %zu\n", sizeof(kt_grad_opts), sizeof(kt_grad_stats), offsetof(kt_grad_stats, chunks));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    return [int(v) for v in out]


def test_ctypes_layout_matches_header():
    v = _c_struct_sizes()
    assert v[:6] == [C.sizeof(_abi.KtConfig), C.sizeof(_abi.KtCamera), C.sizeof(_abi.KtDerived),
                     C.sizeof(_abi.KtTraceStats), C.sizeof(_abi.KtLaunchOpts), C.sizeof(_abi.KtFrameView)]
    assert v[6:10] == [_abi.KtConfig.noise_seed.offset, _abi.KtConfig.bloom_threshold.offset,
                       _abi.KtConfig.precision.offset, _abi.KtConfig.rhs_mode.offset]
    assert v[10:] == [C.sizeof(_abi.KtGradOpts), C.sizeof(_abi.KtGradStats), _abi.KtGradStats.chunks.offset]


def test_c_defaults_equal_python_defaults():
    c = _abi.KtConfig()
    check(lib().kt_config_defaults(C.byref(c)))
    assert c.as_dict() == K.default_config().as_dict()


def _validate(cfg):
    d = _abi.KtDerived()
    rc = lib().kt_validate_config(C.byref(cfg), C.byref(d))
    if rc != 0:
        raise KerrTraceError(rc, "")
    return d


@pytest.mark.parametrize("name", ["cpu256_a09", "hd1080_a0998", "grad512"])
def test_host_derivation_matches_oracle(name):
    cfg, _ = K.preset(name)
    d = _validate(cfg)
    od, odisk, obb = O.derive(cfg)
    for f in ("rh", "disk_inner", "disk_outer", "temperature_scale", "disk_lut_max", "bloom_sigma"):
        assert getattr(d, f) == pytest.approx(getattr(od, f), rel=1e-15, abs=0), f
    assert (d.bloom_radius_int, d.bloom_active) == (od.bloom_radius_int, od.bloom_active)
    disk = np.zeros(cfg.lut_size)
    bb = np.zeros(cfg.lut_size * 3)
    check(lib().kt_host_luts(C.byref(cfg), disk.ctypes.data_as(C.POINTER(C.c_double)),
                             bb.ctypes.data_as(C.POINTER(C.c_double))))
    np.testing.assert_allclose(disk, odisk, rtol=1e-14, atol=0)
    np.testing.assert_allclose(bb.reshape(-1, 3), obb, rtol=1e-13, atol=1e-300)


def test_reference_defaults_derivation():
    """config.yaml defaults: spin -0.45, M 0.5 -> r_h, ISCO (utils.rs:73-77),
    bloom radius 48 -> R 48, sigma 16 (init.rs:58-66, mod.rs:141)."""
    cfg = K.default_config()
    d = _validate(cfg)
    assert d.rh == pytest.approx(0.5 + math.sqrt(0.25 - 0.45**2), rel=1e-15)
    assert d.disk_inner == pytest.approx(0.5 * O.isco(0.9), rel=1e-15)  # a/M = -0.9, odd cbrt
    assert d.bloom_radius_int == 48 and d.bloom_sigma == 16.0 and d.bloom_active == 1
    cfg.bloom_radius = 47.2
    assert _validate(cfg).bloom_radius_int == 48
    cfg.bloom_intensity = 0.0
    assert _validate(cfg).bloom_active == 0


@pytest.mark.parametrize("field,value", [
    ("width", 0), ("height", 0), ("mass", 0.0), ("mass", float("nan")), ("spin", float("inf")),
    ("noise_scale", float("nan")), ("lut_size", 1), ("wavelength_step", 0.0), ("bloom_intensity", -1.0),
    ("bloom_radius", -2.0), ("bloom_radius", float("nan")), ("disk_outer_radius", 1.0), ("spp", 0),
    ("max_attempts", 0), ("precision", 7), ("disk_root_mode", 3), ("rhs_mode", 2),
])
def test_invalid_configs_are_rejected(field, value):
    """build_kerr_params / validate_bloom_settings / LUT input checks
    (utils.rs:53-114, 177-192; init.rs:43-73; blackbody.rs:63-75)."""
    cfg = K.default_config()
    setattr(cfg, field, value)
    rc = lib().kt_validate_config(C.byref(cfg), None)
    assert rc == _abi.KT_ERR_INVALID_CONFIG
    buf = C.create_string_buffer(256)
    lib().kt_last_error(None, buf, 256)
    assert buf.value
    if field not in ("precision", "disk_root_mode", "rhs_mode"):  # product-only knobs
        with pytest.raises(O.OracleError):
            O.derive(cfg)


def test_camera_helpers_match_python():
    c1 = _abi.KtCamera()
    check(lib().kt_camera_observer(100.0, 85.0, 20.0, C.byref(c1)))
    c2 = K.observer_camera(100.0, 85.0, 20.0)
    for f in ("pos", "fwd", "rgt", "up"):
        assert list(getattr(c1, f)) == list(getattr(c2, f))
    assert c1.fov_scale == c2.fov_scale
    pos = (C.c_double * 3)(80.0, 6.0, 0.0)
    c3 = _abi.KtCamera()
    check(lib().kt_camera_from_yaw_pitch(pos, -96.0, -4.0, 15.0, C.byref(c3)))
    c4 = K.camera_from_yaw_pitch((80.0, 6.0, 0.0), -96.0, -4.0, 15.0)
    for f in ("fwd", "rgt", "up"):
        np.testing.assert_allclose(list(getattr(c3, f)), list(getattr(c4, f)), rtol=0, atol=1e-15)
    # the closed-form observer camera is the yaw=-90, pitch=i-90 camera of src/math.rs:3-14
    c5 = K.camera_from_yaw_pitch(tuple(c2.pos), -90.0, 85.0 - 90.0, 20.0)
    for f in ("fwd", "rgt", "up"):
        np.testing.assert_allclose(list(getattr(c5, f)), list(getattr(c2, f)), rtol=0, atol=1e-15)
    assert lib().kt_camera_observer(-1.0, 10.0, 20.0, C.byref(c1)) == _abi.KT_ERR_INVALID_ARGUMENT


def test_null_arguments_return_errors():
    L = lib()
    assert L.kt_config_defaults(None) == _abi.KT_ERR_INVALID_ARGUMENT
    assert L.kt_validate_config(None, None) == _abi.KT_ERR_INVALID_ARGUMENT
    assert L.kt_finish_frame(None, 0) == _abi.KT_ERR_INVALID_ARGUMENT
    assert L.kt_trace(None, None, None, None, None) == _abi.KT_ERR_INVALID_ARGUMENT
    assert L.kt_get_trace_stats(None, None) == _abi.KT_ERR_INVALID_ARGUMENT
    L.kt_destroy(None)


def test_buffer_guard_rejects_mismatched_tensors():
    """Caller buffers reach the kernels as raw pointers, so the Python host
    refuses any tensor that is not the exact contiguous device array a
    kernel indexes (CPU tensors stand in here; with a GPU the same guard
    runs on device tensors)."""
    import torch

    from kerrtrace_amd.renderer import _check_tensor
    good = torch.empty((4, 6, 4), dtype=torch.float32)
    for t, shape, dt in [(good, (4, 6, 3), torch.float32), (good, (4, 6, 4), torch.float64),
                         (good.transpose(0, 1), (6, 4, 4), torch.float32), (good, (4, 6, 4), torch.float32),
                         ("not a tensor", (4, 6, 4), torch.float32)]:
        with pytest.raises(ValueError):
            _check_tensor(t, shape, dt)
