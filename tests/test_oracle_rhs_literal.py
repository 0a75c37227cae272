"""Pin of the oracle's right-hand side against the reference's component
form (cuda/kerr_math.cuh:82-156).

The oracle (and the device) evaluate the Hamiltonian right-hand side in a
factored form: d(l.p)/dx_j and dH/dx_j collapse onto dr/dx_j.  The
reference spells every partial derivative out: dD/dx_j, dN/dx_j, dl/dx_j,
dDenH/dx_j, dH/dx_j, d(l.p)/dx_j.  Below that component form is written
out line by line in float64 (rsqrtf -> 1/sqrt, __fdividef -> /), once as
written (dr^2/dy = y (1 + (u + 4a^2) / 2s), kerr_math.cuh:97) and once with
the exact dr^2/dy = y (1 + (u + 2a^2) / s), and compared with the oracle's
two right-hand-side modes by value.  The two forms share no code, so a
slip in a factored coefficient fails here: the mutations 4a^2 -> 2a^2 in
the as-written w2, a dropped HY term, a sign flip in the a p_z term of
dp_x, and 3 a^2 y^2 -> a^2 y^2 in HR were applied to the oracle on the CPU
and each failed this test.
"""
import numpy as np
import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O


def _component_rhs(s, pt, a, M, exact):
    """cuda/kerr_math.cuh:82-156 as written, in float64."""
    x, y, z, spx, spy, spz = s
    a2 = a * a
    rho2 = x * x + (y * y + z * z)
    u = rho2 - a2
    s_sq = u * u + 4.0 * a2 * y * y
    inv_s = 1.0 / np.sqrt(s_sq)
    r2 = 0.5 * (u + s_sq * inv_s)
    inv_r = 1.0 / np.sqrt(r2)
    r = r2 * inv_r
    dr2_dx = x * (1.0 + u * inv_s)
    if exact:
        dr2_dy = y * (1.0 + (u + 2.0 * a2) * inv_s)
    else:
        dr2_dy = y * (1.0 + (u + 4.0 * a2) * (0.5 * inv_s))
    dr2_dz = z * (1.0 + u * inv_s)
    inv_2r = 0.5 * inv_r
    drdx, drdy, drdz = dr2_dx * inv_2r, dr2_dy * inv_2r, dr2_dz * inv_2r
    denom = r2 + a2
    inv_denom = 1.0 / denom
    inv_denom2 = inv_denom * inv_denom
    lx = (r * x - a * z) * inv_denom
    ly = y * inv_r
    lz = (r * z + a * x) * inv_denom
    r3 = r2 * r
    denomH = r2 * r2 + a2 * y * y
    inv_denomH = 1.0 / denomH
    H = M * r3 * inv_denomH
    lp = lx * spx + (ly * spy + (lz * spz - pt))
    chl = 2.0 * H * lp
    dx, dy, dz = spx - chl * lx, spy - chl * ly, spz - chl * lz
    dD = [2.0 * r * drdx, 2.0 * r * drdy, 2.0 * r * drdz]
    Nx = r * x - a * z
    Nz = r * z + a * x
    dNx = [drdx * x + r, drdy * x, drdz * x - a]
    dNz = [drdx * z + a, drdy * z, drdz * z + r]
    dlx = [(dNx[j] * denom - Nx * dD[j]) * inv_denom2 for j in range(3)]
    dlz = [(dNz[j] * denom - Nz * dD[j]) * inv_denom2 for j in range(3)]
    inv_r2 = inv_r * inv_r
    dly = [-y * drdx * inv_r2, (r - y * drdy) * inv_r2, -y * drdz * inv_r2]
    dDenH = [4.0 * r3 * drdx, 4.0 * r3 * drdy + 2.0 * a2 * y, 4.0 * r3 * drdz]
    inv_denomH2 = inv_denomH * inv_denomH
    factor_H = M * inv_denomH2
    dr = [drdx, drdy, drdz]
    dH = [factor_H * (3.0 * r2 * dr[j] * denomH - r3 * dDenH[j]) for j in range(3)]
    dlp = [dlx[j] * spx + dly[j] * spy + dlz[j] * spz for j in range(3)]
    lp2 = lp * lp
    dp = [dH[j] * lp2 + chl * dlp[j] for j in range(3)]
    return np.array([dx, dy, dz, dp[0], dp[1], dp[2]])


def _states(rng, n):
    """Random states across the regimes the tracer meets, all outside the
    horizon (inside it, near the ring singularity, both forms lose digits to
    cancellation in u + s): far field, the photon-ring zone, just outside
    the horizon and close to the equator (|y| tiny), with momenta of either
    sign."""
    out = []
    for k in range(n):
        band = k % 4
        if band == 0:
            pos = rng.uniform(-60, 60, 3)
        elif band == 1:
            d = rng.normal(size=3)
            pos = d / np.linalg.norm(d) * rng.uniform(2.5, 6.0)
        elif band == 2:
            d = rng.normal(size=3)
            pos = d / np.linalg.norm(d) * rng.uniform(1.4, 2.2)
        else:
            ang, rad = rng.uniform(0, 2 * np.pi), rng.uniform(2.0, 10.0)
            pos = np.array([rad * np.cos(ang), rng.uniform(-1e-3, 1e-3), rad * np.sin(ang)])
        out.append((np.concatenate([pos, rng.normal(size=3)]), -rng.uniform(0.5, 2.0)))
    return out


@pytest.mark.parametrize("a,M", [(0.0, 1.0), (0.5, 1.0), (0.998, 1.0), (-0.45, 0.5)])
@pytest.mark.parametrize("mode", ["reference", "hamiltonian"])
def test_oracle_rhs_equals_component_form(a, M, mode):
    cfg, _ = K.preset("cpu256_a09")
    cfg.spin, cfg.mass = a, M
    exact = mode == "hamiltonian"
    cfg.rhs_mode = K.KT_RHS_HAMILTONIAN if exact else K.KT_RHS_REFERENCE
    rng = np.random.default_rng(1234)
    worst = 0.0
    for s, pt in _states(rng, 400):
        got = O.get_derivs(cfg, s, pt)
        want = _component_rhs(s, pt, a, M, exact)
        # relative error; dx/dl = p - c l is measured against |p| as well,
        # since it may cancel to near zero
        scale = np.maximum(np.abs(want), 1e-300)
        scale[:3] = np.maximum(scale[:3], np.abs(s[3:]))
        worst = max(worst, float(np.max(np.abs(got - want) / scale)))
    assert worst <= 1e-11, worst
