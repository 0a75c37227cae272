"""Loose end-to-end check of a rendered frame against the reference's one
artefact, picture/diagram.png (a 2100x900 first frame saved by
src/app.rs:207-237; its summary is tests/golden/diagram_first_frame.json,
written by tests/golden/make_diagram_fixture.py).

The artefact is an FP32 fast-math frame of an unknown commit, so it is no
pixel target.  What is compared, on a 210x90 grid of block means of the
Rec. 709 luma of the 8-bit sRGB pixels:
  * the bright mask (block luma > 60: the disk and its bloom), by IoU;
  * the block-luma correlation (placement of disk, shadow and sky lines);
  * the 16-bin histogram of the per-pixel luma, by L1 distance.
The scene is config.yaml's defaults (spin -0.45, M 0.5, camera (80, 6, 0)
clamped to the escape radius 80, yaw -96, pitch -4, FOV 15 deg, spp 4,
max 100 steps, as-written right-hand side, secant disk root, tolerance
scale floor 0) rendered at 420x180 with the bloom radius scaled by the
same factor 1/5.  The thresholds sit below what the defaults reach and
above what each of these mistakes reaches (measured with the FP64
oracle): spin sign flipped (IoU 0.15), M = 1 (0.01), yaw -90 deg (0.22),
pitch +4 deg (0.05), FOV 20 deg (0.32, histogram 0.27).
"""
import base64
import json
import os

import numpy as np

from kerrtrace_amd import config as K

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "diagram_first_frame.json")
GRID = (90, 210)             # rows, columns of the comparison grid
RENDER = (420, 180)          # width, height of the check render
BRIGHT = 60.0                # block luma of the bright mask
MIN_BRIGHT_IOU = 0.40
MIN_CORR = 0.70
MAX_HIST_L1 = 0.15


def scene():
    """config.yaml defaults at 420x180 (as-written modes, spp 4)."""
    cfg = K.default_config()
    full_w = cfg.width
    cfg.width, cfg.height = RENDER
    cfg.bloom_radius = cfg.bloom_radius * RENDER[0] / full_w
    cam = K.camera_from_yaw_pitch(K.clamp_to_escape_radius((80.0, 6.0, 0.0), cfg.escape_radius), -96.0, -4.0, 15.0)
    return cfg, cam


def luma_u8(rgb):
    rgb = np.asarray(rgb, dtype=np.float64)
    return 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]


def rgb_of_ldr(ldr):
    v = np.asarray(ldr).astype(np.uint32)
    return np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], axis=-1).astype(np.uint8)


def summary(rgb):
    """(block luma on the GRID rounded to integers, 16-bin luma histogram)."""
    lum = luma_u8(rgb)
    H, W = lum.shape
    fy, fx = H // GRID[0], W // GRID[1]
    blocks = lum[:GRID[0] * fy, :GRID[1] * fx].reshape(GRID[0], fy, GRID[1], fx).mean(axis=(1, 3))
    hist = np.histogram(lum, bins=16, range=(0.0, 256.0))[0] / lum.size
    return np.rint(blocks).astype(np.uint8), hist


def load_fixture():
    with open(FIXTURE) as fh:
        fx = json.load(fh)
    blocks = np.frombuffer(base64.b64decode(fx["block_luma_u8"]), dtype=np.uint8).reshape(GRID)
    return blocks, np.asarray(fx["hist16"], dtype=np.float64)


def metrics(ldr):
    """Compare an LDR frame (uint32 0x00RRGGBB, RENDER size) with the fixture."""
    ref_b, ref_h = load_fixture()
    b, h = summary(rgb_of_ldr(ldr))
    m_ref, m = ref_b > BRIGHT, b > BRIGHT
    iou = float((m_ref & m).sum()) / max(int((m_ref | m).sum()), 1)
    corr = float(np.corrcoef(ref_b.astype(np.float64).ravel(), b.astype(np.float64).ravel())[0, 1])
    return {"bright_iou": iou, "corr": corr, "hist_l1": float(np.abs(ref_h - h).sum()),
            "bright_blocks": int(m.sum()), "bright_blocks_ref": int(m_ref.sum())}


def passes(m):
    return m["bright_iou"] >= MIN_BRIGHT_IOU and m["corr"] >= MIN_CORR and m["hist_l1"] <= MAX_HIST_L1
