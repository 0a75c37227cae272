"""Host-side validation of the differentiable renderer (no GPU needed):
kt_grad_create refuses configurations the adjoint does not cover before it
touches a device."""
import ctypes as C

import pytest

from kerrtrace_amd import config as K
from kerrtrace_amd._abi import (KT_ERR_INVALID_ARGUMENT, KT_ERR_INVALID_CONFIG,
                                KtGradOpts)
from kerrtrace_amd._lib import last_error, lib


def _create(cfg, opts):
    h = C.c_void_p()
    return lib().kt_grad_create(C.byref(cfg), C.byref(opts), C.byref(h)), h


def _ok_opts():
    return KtGradOpts(K.GRAD_FOV_DEG, 32, 512)


@pytest.mark.parametrize("field,value,needle", [
    ("spp", 2, "spp"),
    ("rhs_mode", K.KT_RHS_REFERENCE, "Hamiltonian"),
])
def test_grad_create_rejects_uncovered_configs(field, value, needle):
    cfg, _ = K.preset("grad512")
    setattr(cfg, field, value)
    rc, h = _create(cfg, _ok_opts())
    assert rc == KT_ERR_INVALID_CONFIG and not h.value
    assert needle in last_error(None)


@pytest.mark.parametrize("opts", [KtGradOpts(0.0, 32, 512), KtGradOpts(30.0, 0, 512), KtGradOpts(30.0, 300, 512),
                                  KtGradOpts(30.0, 32, 0), KtGradOpts(30.0, 32, 512, 2), KtGradOpts(30.0, 32, 512, -1)])
def test_grad_create_rejects_bad_options(opts):
    cfg, _ = K.preset("grad512")
    rc, h = _create(cfg, opts)
    assert rc == KT_ERR_INVALID_ARGUMENT and not h.value


def test_grad_null_arguments():
    L = lib()
    assert L.kt_grad_create(None, None, None) == KT_ERR_INVALID_ARGUMENT
    assert L.kt_grad_forward(None, None, None, None, None) == KT_ERR_INVALID_ARGUMENT
    assert L.kt_grad_backward(None, None, None, None, None) == KT_ERR_INVALID_ARGUMENT
    L.kt_grad_destroy(None)
