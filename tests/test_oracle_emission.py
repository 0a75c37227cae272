"""Pins of the oracle's disk emission (cuda/kerr_trace.cuh:174-225) against
what the mathematics and the already-pinned building blocks fix.

The oracle returns, per equatorial crossing of a ray, the pieces its
shading assembled (kto_trace_dir_x): the hit state on the dense output,
the redshift g, the disk-table coordinate and value, the temperature
before and after the noise modulation, the blackbody colour, alpha and
the transmittance before and after.  Each test recomputes one stage of the
assembly from the previous stage's recorded inputs with independent code
(closed forms of the equatorial Kerr-Schild geometry, the table fetches and
fbm that other tests pin) and checks the oracle's recorded output.  Each
docstring names a plausible mistake the test fails; the mutations were
applied to the oracle on the CPU and the test failed for each.
"""
import math

import numpy as np
import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O

F = {name: i for i, name in enumerate(O.CROSS_FIELDS)}


def _scene(a, noise=True, mdot=1.0, incl=80.0, tol=1e-10):
    cfg, cam = K.preset("cpu256_a09")
    cfg.spin = a
    cfg.tolerance = tol
    cfg.mdot_scale = mdot
    cfg.noise_enabled = 1 if noise else 0
    cam = K.observer_camera(50.0, incl, 30.0)
    return cfg, cam


def _rays(cfg, cam, n=10, stride=None):
    """Trace an n x n grid of the frame's pixels; yield (direction, col,
    class, crossing records, (pt, kt))."""
    step = stride or cfg.width // n
    for j in range(n):
        for i in range(n):
            x, y = i * step + step // 2, j * step + step // 2
            d = O.pixel_dir(cfg, cam, x, y, 0)
            col, cls, rec, pk = O.trace_dir_crossings(cfg, list(cam.pos), d)
            yield d, col, cls, rec, pk


def _hits(cfg, cam):
    """All crossing records of the grid that reached the shading."""
    out = []
    for d, col, cls, rec, pk in _rays(cfg, cam):
        for row in rec:
            if row[F["orbit"]] == 1.0:
                out.append((row, pk, d))
    return out


def _redshift_closed_form(row, pk, a, M, Lz):
    """Redshift of an equatorial hit from the Keplerian emitter and the
    photon's conserved quantities: g = sqrt(1 - 3M/r + 2a sqrt(M)/r^1.5) /
    (1 - Omega L/E), Omega = -sqrt(M)/(r^1.5 + a sqrt(M)) (spin along +y),
    E = -p_t, and L = Lz + frame term.  On the equator (y = 0) the
    Kerr-Schild radius is r^2 = x^2 + z^2 - a^2, H = M/r, and the reference's
    l . xi = (l_x z - l_z x) reduces to -a (x^2 + z^2 = r^2 + a^2), so the
    reference's frame term 2H (l . xi)(l . p + kt) is -2a (M/r)(l . p + kt)."""
    x, z = row[F["x"]], row[F["z"]]
    px, pz = row[F["px"]], row[F["pz"]]
    pt, kt = pk
    r = math.sqrt(x * x + z * z - a * a)
    lx = (r * x - a * z) / (r * r + a * a)
    lz = (r * z + a * x) / (r * r + a * a)
    lp = lx * px + lz * pz
    L = Lz - 2.0 * a * (M / r) * (lp + kt)
    disc = 1.0 - 3.0 * M / r + 2.0 * a * math.sqrt(M) / r**1.5
    omega = -math.sqrt(M) / (r**1.5 + a * math.sqrt(M))
    return math.sqrt(disc) / (1.0 - omega * L / (-pt)), r


@pytest.mark.parametrize("a", [0.5, 0.9, 0.998, -0.45])
def test_redshift_frame_term_and_axial_symmetry(a):
    """a != 0 redshift (cuda/kerr_trace.cuh:177-195).  Two facts pin it:
    (1) Lz = z p_x - x p_z generates rotations about the spin axis (y), a
    symmetry of the Kerr-Schild Hamiltonian, so its value at the hit (on
    the dense output) equals its value at the camera; (2) g equals the
    closed form above built from the camera's Lz.  Mutations that fail:
    the frame term's sign flipped or its factor 2 dropped (g off by up to
    30%), l . xi written l_z x - l_x z, x and z swapped in Lz (the
    conservation check), the emitter's Omega with +sqrt(M)."""
    cfg, cam = _scene(a, noise=False)
    hits = _hits(cfg, cam)
    assert len(hits) >= 40
    M = cfg.mass
    worst_lz = worst_g = 0.0
    for row, pk, d in hits:
        s0, _, _ = O.init_ray(cfg, list(cam.pos), d)
        Lz_cam = s0[2] * s0[3] - s0[0] * s0[5]
        Lz_hit = row[F["z"]] * row[F["px"]] - row[F["x"]] * row[F["pz"]]
        worst_lz = max(worst_lz, abs(Lz_hit - Lz_cam) / max(abs(Lz_cam), 1.0))
        g, r = _redshift_closed_form(row, pk, a, M, Lz_cam)
        assert abs(r - row[F["r"]]) <= 1e-12 * r
        worst_g = max(worst_g, abs(row[F["g"]] - g) / g)
    assert worst_lz <= 1e-7, worst_lz
    assert worst_g <= 1e-7, worst_g


@pytest.mark.parametrize("a", [0.9, -0.45])
def test_temperature_without_noise_is_table_times_redshift(a):
    """Noise strength 0: T = T_scale * LUT(u) * g with u = (r - r_in) /
    (r_out - r_in) (cuda/kerr_trace.cuh:196-199): T_scale from the host
    derivation (peak normalisation pinned elsewhere), LUT(u) from the
    texture fetch (pinned against the clamp-address linear filter), g from
    the closed form of test_redshift_frame_term_and_axial_symmetry.
    Mutations that fail: u measured from 0 instead of r_in, the table read
    at r instead of u, g applied as 1/g or squared."""
    cfg, cam = _scene(a, noise=False)
    der, _, _ = O.derive(cfg)
    hits = _hits(cfg, cam)
    assert len(hits) >= 40
    for row, pk, d in hits:
        assert np.isnan(row[F["modulator"]])
        s0, _, _ = O.init_ray(cfg, list(cam.pos), d)
        Lz = s0[2] * s0[3] - s0[0] * s0[5]
        g, r = _redshift_closed_form(row, pk, a, cfg.mass, Lz)
        u = (row[F["r"]] - der.disk_inner) / (der.disk_outer - der.disk_inner)
        assert row[F["u"]] == pytest.approx(u, rel=1e-15, abs=1e-16)
        lut = O.fetch_disk(cfg, u)
        assert row[F["lut"]] == lut
        T = der.temperature_scale * lut * g
        assert row[F["T_base"]] == pytest.approx(T, rel=1e-7)
        assert row[F["T"]] == row[F["T_base"]]


def test_temperature_scales_as_quarter_power_of_accretion_rate():
    """T scales as mdot^(1/4) (Novikov-Thorne T ~ F^(1/4), F ~ mdot): at
    mdot = 1/16 every crossing's temperature is exactly half the mdot = 1
    value, before and after the noise, on identical trajectories (a dimmer
    ray may cross more often before its transmittance runs out, so the
    common crossings are compared).  Fails if mdot enters as mdot^(1/2) or
    mdot, or also changes the trajectory."""
    cfg1, cam = _scene(0.9, noise=True)
    cfg2, _ = _scene(0.9, noise=True, mdot=1.0 / 16.0)
    n = 0
    for (d, _, _, r1, _), (_, _, _, r2, _) in zip(_rays(cfg1, cam, n=6), _rays(cfg2, cam, n=6)):
        assert len(r2) >= len(r1)
        for a1, a2 in zip(r1, r2):
            np.testing.assert_array_equal(a1[:F["u"] + 2], a2[:F["u"] + 2])
            if a1[F["orbit"]] == 1.0:
                assert a2[F["T_base"]] == 0.5 * a1[F["T_base"]]
                assert a2[F["T"]] == 0.5 * a1[F["T"]]
                n += 1
    assert n >= 10


def _bl_phi_equator(x, z, a):
    """Boyer-Lindquist azimuth of an equatorial Kerr-Schild point: from
    x = r sin(phi) + a cos(phi), z = r cos(phi) - a sin(phi) (sin theta = 1),
    cos(phi) = (r z + a x)/(r^2 + a^2), sin(phi) = (r x - a z)/(r^2 + a^2)."""
    r = math.sqrt(x * x + z * z - a * a)
    return math.atan2((r * x - a * z) / (r * r + a * a), (r * z + a * x) / (r * r + a * a)), r


@pytest.mark.parametrize("a,octaves,seed", [(0.9, 4, 0), (0.998, 6, 1), (-0.45, 10, 0)])
def test_noise_modulation_closed_form(a, octaves, seed):
    """Noise on (cuda/kerr_trace.cuh:200-214): T / T_base = 1 + s (fbm(nu,
    nu') - 1/2) with the spiral phase phi + w Omega, Omega = 1/(a +
    sqrt(r^3/M)), (nu, nu') = r scale (cos, sin)(phase), phi the BL azimuth
    of the hit (closed form above), fbm pinned by its own tests.  Mutations
    that fail: the winding added with the wrong sign, Omega with a - sqrt,
    nu and nu' swapped, the KS azimuth atan2(x, z) used for phi, the
    modulator applied as s (fbm + 1/2)."""
    cfg, cam = _scene(a, noise=True)
    cfg.noise_detail = octaves
    cfg.noise_seed = seed
    hits = _hits(cfg, cam)
    assert len(hits) >= 40
    s, w, sc, M = cfg.noise_strength, cfg.noise_winding, cfg.noise_scale, cfg.mass
    for row, pk, d in hits:
        phi, r = _bl_phi_equator(row[F["x"]], row[F["z"]], a)
        assert row[F["phi"]] == pytest.approx(phi, abs=1e-13)
        assert row[F["r_bl"]] == pytest.approx(r, rel=1e-13)
        omega = 1.0 / (a + math.sqrt(r**3 / M))
        phase = phi + w * omega
        nu, nv = r * math.cos(phase) * sc, r * math.sin(phase) * sc
        mod = 1.0 + s * (O.fbm(nu, nv, octaves, seed) - 0.5)
        assert row[F["modulator"]] == pytest.approx(mod, rel=1e-9)
        assert row[F["T"]] == pytest.approx(row[F["T_base"]] * mod, rel=1e-15)


@pytest.mark.parametrize("a,noise", [(0.9, True), (0.998, False), (0.5, True)])
def test_colour_alpha_and_front_to_back_compositing(a, noise):
    """Blackbody colour, opacity and compositing (cuda/kerr_trace.cuh:215-225):
    rgb_k = the blackbody fetch at the crossing's final T (fetch pinned
    against the texture-filter definition); alpha_k = 1 - exp(-luma_k /
    (2 dl)) with Rec. 709 luma; the transmittance before crossing k is
    prod_{j<k} (1 - alpha_j); the ray's colour is sum_k rgb_k alpha_k Tr_k
    (front to back) on rays that do not end on a sky grid line, whose sky
    term is zero.  Mutations that fail: rgb fetched at T before the noise,
    alpha without the 1/2, luma with the Rec. 601 weights, Tr updated before
    the colour is added (back-to-front order), the colour summed without
    Tr.  A dim disk (peak 4000 K) keeps alpha well below 1, so many rays
    composite two or more crossings."""
    cfg, cam = _scene(a, noise=noise)
    cfg.disk_peak_temperature = 4000.0
    dl = cfg.wavelength_step
    n_multi = n_comp = 0
    for d, col, cls, rec, pk in _rays(cfg, cam, n=12):
        tr = 1.0
        acc = np.zeros(3)
        comp = 0
        for row in rec:
            assert row[F["tr_before"]] == tr
            if row[F["composited"]] == 1.0 or not np.isnan(row[F["alpha"]]):
                rgb, ovf = O.fetch_blackbody(cfg, row[F["T"]])
                assert not ovf
                got = row[[F["r_col"], F["g_col"], F["b_col"]]]
                np.testing.assert_array_equal(got, rgb)
                luma = 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]
                alpha = 1.0 - math.exp(-luma / (2.0 * dl))
                assert row[F["alpha"]] == pytest.approx(alpha, rel=1e-13, abs=1e-300)
                acc += rgb * alpha * tr
                tr = tr * (1.0 - alpha)
                assert row[F["tr_after"]] == pytest.approx(tr, rel=1e-15)
                tr = row[F["tr_after"]]
                comp += 1 if alpha > 0.0 else 0
            else:
                assert row[F["tr_after"]] == tr
        n_comp += comp
        n_multi += comp >= 2
        assert ((cls >> 5) & 7) == min(comp, 7)
        assert bool(cls & 4) == (comp > 0)
        if not cls & 8:
            np.testing.assert_allclose(col, acc, rtol=1e-14, atol=1e-300)
    assert n_comp >= 40 and n_multi >= 2
