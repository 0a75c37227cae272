"""Configuration: the reference's ``config.yaml`` schema, extensions, presets.

``load_config`` reads a YAML file laid out like the reference's
``config.yaml`` (schema: src/main.rs:14-122) into a :class:`KtConfig` plus a
camera.  Integer knobs accept ``1e2``-style values as the reference's
``deserialize_u32_scientific`` does (src/main.rs:123-189).  An optional
``kerrtrace_amd`` section carries the knobs the reference does not have
(precision, disk-root mode, noise seed, bloom threshold, peak temperature,
accretion-rate scale, tolerance floor, observer placement).

Presets name the five workloads the renderer is measured and checked on,
by what they are.
"""
import json
import math
import os

import yaml

from ._abi import (KT_DISK_ROOT_SECANT_REF, KT_DISK_ROOT_TRUE,  # noqa: F401
                   KT_END_ESCAPE, KT_END_HORIZON, KT_END_OTHER, KT_FLAG_DISK,
                   KT_FLAG_SKY_GRID, KT_PRECISION_FP64, KT_PRECISION_MIXED,
                   KT_RHS_HAMILTONIAN, KT_RHS_REFERENCE, KtCamera, KtConfig)

U32_MAX = 0xFFFFFFFF


class ConfigError(ValueError):
    """Invalid configuration (the reference returns anyhow errors)."""


def parse_u32(value, name="value"):
    """u32 from int, integral float or numeric string (src/main.rs:123-189)."""
    if isinstance(value, bool):
        raise ConfigError(f"{name}: a non-negative integer is required")
    if isinstance(value, int):
        if value < 0:
            raise ConfigError(f"{name}: integer must be non-negative")
        if value > U32_MAX:
            raise ConfigError(f"{name}: integer out of range for u32")
        return value
    if isinstance(value, str):
        try:
            value = float(value)
        except ValueError as exc:
            raise ConfigError(f"{name}: failed to parse number") from exc
    if isinstance(value, float):
        if not math.isfinite(value):
            raise ConfigError(f"{name}: number must be finite")
        if value != int(value):
            raise ConfigError(f"{name}: number must be an integer")
        if value < 0.0 or value > float(U32_MAX):
            raise ConfigError(f"{name}: number out of range for u32")
        return int(value)
    raise ConfigError(f"{name}: a non-negative integer is required")


def parse_float(value, name="value"):
    """Float from a YAML scalar; ``1e-5`` arrives as a string in YAML 1.1."""
    if isinstance(value, bool):
        raise ConfigError(f"{name}: a number is required")
    if isinstance(value, (int, float)):
        return float(value)
    if isinstance(value, str):
        try:
            return float(value)
        except ValueError as exc:
            raise ConfigError(f"{name}: failed to parse number") from exc
    raise ConfigError(f"{name}: a number is required")


# Mixed precision: renders whose tolerance is below this trace in FP64 (FP32
# rounding of the right-hand side would reach the error tolerance).
MIXED_MIN_TOLERANCE_DEFAULT = 5e-8


def default_config():
    """The reference's config.yaml defaults (config.yaml:1-67)."""
    c = KtConfig()
    c.width, c.height = 2100, 900
    c.spp = 4
    c.sky_grid_divisions = 40
    c.exposure_scale = 3e-3
    c.sky_line_thickness = 5e-4
    c.sky_intensity = 30.0
    c.spin, c.mass = -0.45, 0.5
    c.disk_outer_radius = 40.0
    c.disk_temperature_scale = 3e4
    c.initial_step = 1.0
    c.tolerance = 1e-5
    c.max_steps = 100
    c.max_attempts = 3
    c.transmittance_cutoff = 1e-2
    c.horizon_epsilon = 5e-7
    c.escape_radius = 80.0
    c.lut_size = 4096
    c.lut_max_temp = 3e4
    c.wavelength_start, c.wavelength_end, c.wavelength_step = 380.0, 720.0, 10.0
    c.noise_enabled = 1
    c.noise_detail = 10
    c.noise_scale, c.noise_strength, c.noise_winding = 0.1, 0.6, 50.0
    c.noise_seed = 0
    c.bloom_enabled = 1
    c.bloom_intensity = 0.5
    c.bloom_radius = 48.0
    c.bloom_threshold = 0.0
    c.tol_floor = 0.0
    c.disk_peak_temperature = 0.0
    c.mdot_scale = 1.0
    c.disk_root_mode = KT_DISK_ROOT_SECANT_REF
    c.precision = KT_PRECISION_FP64
    c.rhs_mode = KT_RHS_REFERENCE
    c.mixed_min_tolerance = MIXED_MIN_TOLERANCE_DEFAULT
    return c


def camera_from_yaw_pitch(pos, yaw_deg, pitch_deg, fov_deg):
    """Camera basis from yaw/pitch (src/math.rs:3-14), fov_scale = tan(fov/2)
    (src/app.rs:163), evaluated in FP64."""
    yaw = math.radians(yaw_deg)
    pitch = math.radians(pitch_deg)
    f = (math.sin(yaw) * math.cos(pitch), math.sin(pitch),
         -math.cos(yaw) * math.cos(pitch))
    f = _normalize(f)
    world_up = (0.0, 1.0, 0.0)
    r = _normalize(_cross(f, world_up))
    u = _normalize(_cross(r, f))
    return _camera(pos, f, r, u, math.tan(math.radians(fov_deg) / 2.0))


def observer_camera(r_obs, inclination_deg, fov_deg):
    """Observer at distance r_obs and inclination i from the spin axis (+y),
    looking at the hole: pos = r (sin i, cos i, 0), fwd = -pos/r,
    rgt = (0, 0, -1), up = (-cos i, sin i, 0).  This is the yaw = -90 deg,
    pitch = i - 90 deg camera of src/math.rs:3-14 written in closed form
    (the closed form avoids cos(-pi/2) = 6e-17 in the basis)."""
    i = math.radians(inclination_deg)
    si, ci = math.sin(i), math.cos(i)
    pos = (r_obs * si, r_obs * ci, 0.0)
    fwd = (-si, -ci, 0.0)
    rgt = (0.0, 0.0, -1.0)
    up = (-ci, si, 0.0)
    return _camera(pos, fwd, rgt, up, math.tan(math.radians(fov_deg) / 2.0))


def _camera(pos, fwd, rgt, up, fov_scale):
    cam = KtCamera()
    for k in range(3):
        cam.pos[k] = float(pos[k])
        cam.fwd[k] = float(fwd[k])
        cam.rgt[k] = float(rgt[k])
        cam.up[k] = float(up[k])
    cam.fov_scale = float(fov_scale)
    return cam


def _normalize(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return (v[0] / n, v[1] / n, v[2] / n)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0])


def clamp_to_escape_radius(pos, escape_radius):
    """src/app.rs:126-139: a camera outside the escape sphere is pulled onto it."""
    if not (math.isfinite(escape_radius) and escape_radius > 0.0):
        raise ConfigError(f"escape_radius invalid: {escape_radius}")
    dist = math.sqrt(sum(p * p for p in pos))
    if dist > escape_radius:
        s = escape_radius / dist
        return tuple(p * s for p in pos)
    return tuple(pos)


def config_from_dict(tree):
    """KtConfig + camera from a parsed config.yaml tree."""
    c = default_config()
    try:
        win = tree["window"]
        c.width = parse_u32(win["width"], "window.width")
        c.height = parse_u32(win["height"], "window.height")
        k = tree["kernel"]
        c.spp = parse_u32(k["spp"], "kernel.spp")
        c.exposure_scale = parse_float(k["exposure_scale"], "kernel.exposure_scale")
        sky = k["sky"]
        c.sky_grid_divisions = parse_u32(sky["grid_divisions"], "kernel.sky.grid_divisions")
        c.sky_line_thickness = parse_float(sky["line_thickness"], "kernel.sky.line_thickness")
        c.sky_intensity = parse_float(sky["intensity"], "kernel.sky.intensity")
        bh = k["black_hole"]
        c.spin = parse_float(bh["spin"], "kernel.black_hole.spin")
        c.mass = parse_float(bh["mass"], "kernel.black_hole.mass")
        disk = k["disk"]
        c.disk_outer_radius = parse_float(disk["outer_radius"], "kernel.disk.outer_radius")
        c.disk_temperature_scale = parse_float(disk["temperature_scale"], "kernel.disk.temperature_scale")
        it = k["integrator"]
        c.initial_step = parse_float(it["initial_step"], "integrator.initial_step")
        c.tolerance = parse_float(it["tolerance"], "integrator.tolerance")
        c.max_steps = parse_u32(it["max_steps"], "integrator.max_steps")
        c.max_attempts = parse_u32(it["max_attempts"], "integrator.max_attempts")
        c.transmittance_cutoff = parse_float(it["transmittance_cutoff"], "integrator.transmittance_cutoff")
        c.horizon_epsilon = parse_float(it["horizon_epsilon"], "integrator.horizon_epsilon")
        c.escape_radius = parse_float(it["escape_radius"], "integrator.escape_radius")
        bb = tree["blackbody"]
        c.lut_size = parse_u32(bb["lut_size"], "blackbody.lut_size")
        c.lut_max_temp = parse_float(bb["lut_max_temp"], "blackbody.lut_max_temp")
        c.wavelength_start = parse_float(bb["wavelength_start"], "blackbody.wavelength_start")
        c.wavelength_end = parse_float(bb["wavelength_end"], "blackbody.wavelength_end")
        c.wavelength_step = parse_float(bb["wavelength_step"], "blackbody.wavelength_step")
        dn = tree["disk_noise"]
        c.noise_enabled = 1 if bool(dn["enabled"]) else 0
        c.noise_scale = parse_float(dn["scale"], "disk_noise.scale")
        c.noise_strength = parse_float(dn["strength"], "disk_noise.strength")
        c.noise_winding = parse_float(dn["winding"], "disk_noise.winding")
        c.noise_detail = parse_u32(dn["detail"], "disk_noise.detail")
        bl = tree["bloom"]
        c.bloom_enabled = 1 if bool(bl["enabled"]) else 0
        c.bloom_intensity = parse_float(bl["intensity"], "bloom.intensity")
        c.bloom_radius = parse_float(bl["radius"], "bloom.radius")
        cam_t = tree["camera"]
        pos = [parse_float(v, "camera.position") for v in cam_t["position"]]
        yaw = parse_float(cam_t["yaw"], "camera.yaw")
        pitch = parse_float(cam_t["pitch"], "camera.pitch")
        fov = parse_float(cam_t["fov"], "camera.fov")
    except KeyError as exc:
        raise ConfigError(f"missing config key: {exc}") from exc
    ext = tree.get("kerrtrace_amd") or {}
    prec = str(ext.get("precision", "fp64")).lower()
    if prec not in ("fp64", "mixed"):
        raise ConfigError(f"kerrtrace_amd.precision must be fp64 or mixed, got {prec}")
    c.precision = KT_PRECISION_MIXED if prec == "mixed" else KT_PRECISION_FP64
    root = str(ext.get("disk_root", "secant_ref")).lower()
    if root not in ("secant_ref", "true_root"):
        raise ConfigError(f"kerrtrace_amd.disk_root must be secant_ref or true_root, got {root}")
    c.disk_root_mode = KT_DISK_ROOT_TRUE if root == "true_root" else KT_DISK_ROOT_SECANT_REF
    rhs = str(ext.get("rhs", "reference")).lower()
    if rhs not in ("reference", "hamiltonian"):
        raise ConfigError(f"kerrtrace_amd.rhs must be reference or hamiltonian, got {rhs}")
    c.rhs_mode = KT_RHS_HAMILTONIAN if rhs == "hamiltonian" else KT_RHS_REFERENCE
    c.noise_seed = int(ext.get("noise_seed", 0))
    c.bloom_threshold = parse_float(ext.get("bloom_threshold", 0.0), "bloom_threshold")
    c.tol_floor = parse_float(ext.get("tol_floor", 0.0), "tol_floor")
    c.disk_peak_temperature = parse_float(ext.get("peak_temperature", 0.0), "peak_temperature")
    c.mdot_scale = parse_float(ext.get("mdot_scale", 1.0), "mdot_scale")
    c.mixed_min_tolerance = parse_float(ext.get("mixed_min_tolerance", MIXED_MIN_TOLERANCE_DEFAULT), "mixed_min_tolerance")
    obs = ext.get("observer")
    if obs is not None:
        cam = observer_camera(parse_float(obs["r"], "observer.r"),
                              parse_float(obs["inclination"], "observer.inclination"), fov)
    else:
        pos = clamp_to_escape_radius(pos, c.escape_radius)
        cam = camera_from_yaw_pitch(pos, yaw, pitch, fov)
    return c, cam


def load_config(path):
    """Parse a config.yaml laid out like the reference's (src/main.rs:190-194)."""
    with open(path, "r", encoding="utf-8") as fh:
        tree = yaml.safe_load(fh)
    if not isinstance(tree, dict):
        raise ConfigError("config file must hold a mapping")
    return config_from_dict(tree)


# --------------------------------------------------------------------------
# Presets: the measured and checked workloads, named by what they are.
# Knobs a workload does not state take the reference defaults, except:
# spp = 1, tol_floor = 1 (tolerance never tighter than stated), disk root
# solved on the interpolant, exact Hamiltonian right-hand side, escape
# radius = observer distance (the camera sits on the escape sphere as
# src/app.rs:126-139 enforces), bloom radius = 3 sigma.

def _observer_preset(width, height, spin, incl, r_obs, fov, tol, r_out,
                     octaves, seed, sigma, threshold=1.0, peak_t=1e4,
                     max_steps=20000, precision=KT_PRECISION_FP64, mdot=1.0):
    c = default_config()
    c.width, c.height = width, height
    c.spp = 1
    c.spin, c.mass = spin, 1.0
    c.tolerance = tol
    c.max_steps = max_steps
    c.escape_radius = r_obs
    c.disk_outer_radius = r_out
    c.disk_peak_temperature = peak_t
    c.noise_detail = octaves
    c.noise_seed = seed
    c.bloom_radius = 3.0 * sigma
    c.bloom_threshold = threshold
    c.tol_floor = 1.0
    c.disk_root_mode = KT_DISK_ROOT_TRUE
    c.rhs_mode = KT_RHS_HAMILTONIAN
    c.precision = precision
    c.mdot_scale = mdot
    return c, observer_camera(r_obs, incl, fov)


# The differentiable scene: theta = (spin, inclination [rad], r_obs [M],
# accretion-rate scale); the fit starts at GRAD_START and the target frame
# is rendered at GRAD_TARGET.
GRAD_FOV_DEG = 30.0
GRAD_TARGET = (0.7, math.radians(70.0), 60.0, 1.0)
GRAD_START = (0.5, math.radians(55.0), 70.0, 0.8)


def preset(name, **overrides):
    """Return (KtConfig, KtCamera) of a named workload.

    cpu256_a09      256x256, a=0.9, r=50M, i=80 deg, FOV 30 deg, tol 1e-8,
                    disk ISCO..20M, peak 1e4 K, 4 octaves seed 0, sigma 8 px
    hd1080_a0998    1920x1080, a=0.998, r=100M, i=85 deg, FOV 20 deg,
                    tol 1e-9, disk ISCO..30M, 6 octaves seed 1, sigma 12 px
    uhd2160_i60     3840x2160, i=60 deg, tol 1e-7, mixed precision, other
                    knobs as hd1080_a0998 (pass spin=...)
    grad512         512x512 differentiable scene, a=0.7, i=70 deg, r=60M,
                    mdot 1.0, FOV 30 deg, tol 1e-7, FP64, other knobs as
                    cpu256_a09 (theta = GRAD_TARGET)
    """
    if name == "cpu256_a09":
        c, cam = _observer_preset(256, 256, 0.9, 80.0, 50.0, 30.0, 1e-8, 20.0, 4, 0, 8.0)
    elif name == "hd1080_a0998":
        c, cam = _observer_preset(1920, 1080, 0.998, 85.0, 100.0, 20.0, 1e-9, 30.0, 6, 1, 12.0)
    elif name == "uhd2160_i60":
        spin = overrides.pop("spin", 0.998)
        c, cam = _observer_preset(3840, 2160, spin, 60.0, 100.0, 20.0, 1e-7, 30.0, 6, 1, 12.0,
                                  precision=KT_PRECISION_MIXED)
    elif name == "grad512":
        c, cam = _observer_preset(512, 512, overrides.pop("spin", 0.7), overrides.pop("inclination", 70.0),
                                  overrides.pop("r_obs", 60.0), GRAD_FOV_DEG, 1e-7, 20.0, 4, 0, 8.0,
                                  precision=KT_PRECISION_FP64, mdot=overrides.pop("mdot", 1.0))
    else:
        raise KeyError(f"unknown preset {name!r}")
    for key, value in overrides.items():
        if key == "size":
            c.width, c.height = value
        else:
            setattr(c, key, value)
    return c, cam


def scaled(cfg, cam, width, height):
    """Same scene at another resolution (pixel pitch follows the size)."""
    c = cfg.copy()
    c.width, c.height = int(width), int(height)
    return c, cam.copy()


BATCH64_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "presets", "batch64_seed7.json")


def batch_preset(path=BATCH64_PATH, limit=None, max_side=None):
    """The 64-render batch (seed 7 draw, committed JSON): a list of
    (KtConfig, KtCamera).  Knobs the draw does not set follow hd1080_a0998
    (FOV 20 deg, tol 1e-9, disk ISCO..30 M, 6 octaves seed 1, sigma 12 px).
    limit keeps the first renders; max_side caps a render's side (the scene
    is unchanged, only the pixel count)."""
    with open(path, "r", encoding="utf-8") as fh:
        doc = json.load(fh)
    out = []
    for r in doc["renders"][:limit]:
        side = int(r["side"]) if max_side is None else min(int(r["side"]), int(max_side))
        out.append(_observer_preset(side, side, float(r["spin"]), float(r["inclination_deg"]), float(r["r_obs"]),
                                    20.0, 1e-9, 30.0, 6, 1, 12.0))
    return out
