"""config.yaml loading (src/main.rs:14-194): the reference's schema and
defaults, its u32 parsing with scientific notation (src/main.rs:123-189),
the camera clamp onto the escape sphere (src/app.rs:126-139), the
extension section, errors, and a dict -> YAML -> config round trip."""
import math

import pytest
import yaml

from kerrtrace_amd import config as K

# The reference's default configuration written in its schema (the values
# of config.yaml:1-67 that the renderer reads; window/UI-only keys included
# so the loader is shown to ignore them).
DEFAULTS_YAML = """
window: {width: 2100, height: 900, vsync: false}
camera:
  position: [80.0, 6.0, 0.0]
  yaw: -96.0
  pitch: -4.0
  fov: 15.0
  pitch_limit: [-80.0, 80.0]
renderer: {block_dim: [16, 16], save_first_frame: true}
bloom: {enabled: true, intensity: 0.5, radius: 48.0}
blackbody: {lut_size: 4096, lut_max_temp: 3e4, wavelength_start: 380.0, wavelength_end: 720.0,
            wavelength_step: 10.0}
kernel:
  spp: 4
  exposure_scale: 3e-3
  sky: {grid_divisions: 40, line_thickness: 5e-4, intensity: 30.0}
  black_hole: {spin: -0.45, mass: 0.5}
  disk: {outer_radius: 40.0, temperature_scale: 3e4}
  integrator: {initial_step: 1e0, tolerance: 1e-5, max_steps: 1e2, max_attempts: 3,
               transmittance_cutoff: 1e-2, horizon_epsilon: 5e-7, escape_radius: 80.0}
disk_noise: {enabled: true, scale: 0.1, strength: 0.6, winding: 50.0, detail: 10}
"""


def _fields(cfg):
    return {name: getattr(cfg, name) for name, _ in cfg._fields_}


def test_reference_defaults_load_to_default_config(tmp_path):
    p = tmp_path / "config.yaml"
    p.write_text(DEFAULTS_YAML)
    cfg, cam = K.load_config(str(p))
    assert _fields(cfg) == _fields(K.default_config())
    # the camera at (80, 6, 0) lies outside the escape sphere r = 80 and is
    # pulled onto it before the yaw/pitch basis is built
    pos = K.clamp_to_escape_radius((80.0, 6.0, 0.0), 80.0)
    assert math.isclose(math.sqrt(sum(v * v for v in pos)), 80.0, rel_tol=1e-15)
    ref = K.camera_from_yaw_pitch(pos, -96.0, -4.0, 15.0)
    for k in range(3):
        assert cam.pos[k] == ref.pos[k] and cam.fwd[k] == ref.fwd[k]
        assert cam.rgt[k] == ref.rgt[k] and cam.up[k] == ref.up[k]
    assert cam.fov_scale == math.tan(math.radians(7.5))


@pytest.mark.parametrize("value,expect", [(7, 7), ("1e2", 100), (2.0e4, 20000), ("4096", 4096), (0, 0),
                                          (4294967295, 4294967295), ("4.294967295e9", 4294967295)])
def test_u32_scientific_notation(value, expect):
    assert K.parse_u32(value) == expect


@pytest.mark.parametrize("value", [-1, 1.5, "2.5e0", "abc", True, 4294967296, "1e10", float("inf"), float("nan"),
                                   None, [3]])
def test_u32_rejects(value):
    with pytest.raises(K.ConfigError):
        K.parse_u32(value)


def test_floats_in_yaml_11_strings():
    # YAML 1.1 reads 1e-5 (no dot) as a string; the loader takes it as a number
    tree = yaml.safe_load("x: 1e-5")
    assert isinstance(tree["x"], str) and K.parse_float(tree["x"]) == 1e-5
    with pytest.raises(K.ConfigError):
        K.parse_float("fast")
    with pytest.raises(K.ConfigError):
        K.parse_float(True)


def test_extension_section():
    tree = yaml.safe_load(DEFAULTS_YAML)
    tree["kerrtrace_amd"] = {"precision": "mixed", "disk_root": "true_root", "rhs": "hamiltonian",
                             "noise_seed": 7, "bloom_threshold": 1.0, "tol_floor": 1.0,
                             "peak_temperature": 1e4, "mdot_scale": 0.8,
                             "observer": {"r": 60.0, "inclination": 70.0}}
    cfg, cam = K.config_from_dict(tree)
    assert cfg.precision == K.KT_PRECISION_MIXED
    assert cfg.disk_root_mode == K.KT_DISK_ROOT_TRUE
    assert cfg.rhs_mode == K.KT_RHS_HAMILTONIAN
    assert (cfg.noise_seed, cfg.bloom_threshold, cfg.tol_floor) == (7, 1.0, 1.0)
    assert (cfg.disk_peak_temperature, cfg.mdot_scale) == (1e4, 0.8)
    obs = K.observer_camera(60.0, 70.0, 15.0)
    assert [cam.pos[k] for k in range(3)] == [obs.pos[k] for k in range(3)]
    for bad in ({"precision": "fp16"}, {"disk_root": "newton"}, {"rhs": "approx"}):
        t = yaml.safe_load(DEFAULTS_YAML)
        t["kerrtrace_amd"] = bad
        with pytest.raises(K.ConfigError):
            K.config_from_dict(t)


def test_missing_key_names_it():
    tree = yaml.safe_load(DEFAULTS_YAML)
    del tree["kernel"]["integrator"]["tolerance"]
    with pytest.raises(K.ConfigError, match="tolerance"):
        K.config_from_dict(tree)
    with pytest.raises(K.ConfigError):
        K.config_from_dict({"window": {"width": 10}})


def test_non_mapping_file_rejected(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("- 1\n- 2\n")
    with pytest.raises(K.ConfigError):
        K.load_config(str(p))


def _tree_of(cfg, cam_pos, yaw, pitch, fov):
    """The reference-schema tree of a config (the inverse of config_from_dict
    for the fields the schema holds)."""
    return {
        "window": {"width": cfg.width, "height": cfg.height},
        "camera": {"position": list(cam_pos), "yaw": yaw, "pitch": pitch, "fov": fov},
        "bloom": {"enabled": bool(cfg.bloom_enabled), "intensity": cfg.bloom_intensity, "radius": cfg.bloom_radius},
        "blackbody": {"lut_size": cfg.lut_size, "lut_max_temp": cfg.lut_max_temp,
                      "wavelength_start": cfg.wavelength_start, "wavelength_end": cfg.wavelength_end,
                      "wavelength_step": cfg.wavelength_step},
        "kernel": {"spp": cfg.spp, "exposure_scale": cfg.exposure_scale,
                   "sky": {"grid_divisions": cfg.sky_grid_divisions, "line_thickness": cfg.sky_line_thickness,
                           "intensity": cfg.sky_intensity},
                   "black_hole": {"spin": cfg.spin, "mass": cfg.mass},
                   "disk": {"outer_radius": cfg.disk_outer_radius, "temperature_scale": cfg.disk_temperature_scale},
                   "integrator": {"initial_step": cfg.initial_step, "tolerance": cfg.tolerance,
                                  "max_steps": cfg.max_steps, "max_attempts": cfg.max_attempts,
                                  "transmittance_cutoff": cfg.transmittance_cutoff,
                                  "horizon_epsilon": cfg.horizon_epsilon, "escape_radius": cfg.escape_radius}},
        "disk_noise": {"enabled": bool(cfg.noise_enabled), "scale": cfg.noise_scale, "strength": cfg.noise_strength,
                       "winding": cfg.noise_winding, "detail": cfg.noise_detail},
    }


def test_round_trip_through_yaml(tmp_path):
    cfg = K.default_config()
    cfg.width, cfg.height, cfg.spp = 640, 360, 2
    cfg.spin, cfg.mass, cfg.tolerance, cfg.max_steps = 0.93, 1.0, 3e-9, 20000
    cfg.noise_detail, cfg.bloom_radius, cfg.escape_radius = 6, 36.0, 100.0
    p = tmp_path / "rt.yaml"
    p.write_text(yaml.safe_dump(_tree_of(cfg, (10.0, 20.0, -5.0), -80.0, 3.0, 25.0)))
    back, cam = K.load_config(str(p))
    assert _fields(back) == _fields(cfg)
    ref = K.camera_from_yaw_pitch((10.0, 20.0, -5.0), -80.0, 3.0, 25.0)  # inside the sphere: not clamped
    assert [cam.pos[k] for k in range(3)] == [10.0, 20.0, -5.0]
    assert [cam.fwd[k] for k in range(3)] == [ref.fwd[k] for k in range(3)]
