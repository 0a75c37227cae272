"""Writes cpu256_a09_64x64.npz: a golden frame and step logs of the FP64
oracle, so that a later change to the oracle (or to the device tracer that
mirrors it) shows up against fixed numbers, not only against the live
oracle.

Workload: the 256x256 CPU configuration (spin 0.9, M = 1, observer at
r = 50 M, inclination 80 deg, FOV 30 deg, DP5(4) at tolerance 1e-8, max
20000 steps, Novikov-Thorne disk ISCO..20 M at peak 1e4 K, FBM 4 octaves
seed 0) scaled to 64x64 pixels, spp 1.

Arrays:
  hdr        float64 [64, 64, 3]  exposure-scaled linear rgb
  cls        uint8   [64, 64]     class bytes (kerrtrace.h KT_END_*/KT_FLAG_*)
  attempts   uint32  [64, 64]     DP5(4) attempts per pixel
  log_pixels int32   [16, 2]      (x, y) of the logged rays
  log_len    int32   [16]         committed steps of each logged ray
  logs       float64 [sum, 8]     per committed step: h, accepted, x, y, z,
                                   px, py, pz after the step (rays in order)
  log_stats  uint32  [16, 5]      attempts, rhs, accepted, crossings, steps
The 16 rays: the frame's longest ray, then rays of each end class (disk,
horizon, escape, sky grid line) spread over the frame, chosen by a fixed
rule from the class map.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from kerrtrace_amd import config as K  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "cpu256_a09_64x64.npz")
N_LOGS = 16


def scene():
    cfg, cam = K.preset("cpu256_a09")
    return K.scaled(cfg, cam, 64, 64)


def pick_pixels(cls, att):
    """The longest ray, then up to 4 rays of each of the classes disk,
    horizon, escape (off grid), escape on a grid line, evenly spaced in
    raster order, topped up with the longest remaining rays."""
    H, W = cls.shape
    flat_att = att.reshape(-1)
    c = cls.reshape(-1).astype(np.int32)
    picked = [int(np.argmax(flat_att))]
    groups = [(c & 4) != 0, ((c & 3) == 1) & ((c & 4) == 0), ((c & 3) == 2) & ((c & 12) == 0),
              ((c & 3) == 2) & ((c & 8) != 0)]
    for g in groups:
        idx = np.nonzero(g)[0]
        for k in np.linspace(0, len(idx) - 1, num=min(4, len(idx))).astype(int) if len(idx) else []:
            if int(idx[k]) not in picked and len(picked) < N_LOGS:
                picked.append(int(idx[k]))
    for i in np.argsort(-flat_att, kind="stable"):
        if len(picked) >= N_LOGS:
            break
        if int(i) not in picked:
            picked.append(int(i))
    return [(i % W, i // W) for i in picked[:N_LOGS]]


def make():
    cfg, cam = scene()
    hdr, cls, att, _ = O.render(cfg, cam, threads=0)
    pixels = pick_pixels(cls, att)
    logs, lens, stats = [], [], []
    for x, y in pixels:
        d = O.pixel_dir(cfg, cam, x, y, 0)
        col, c, st, log = O.trace_dir(cfg, list(cam.pos), d, max_log=int(cfg.max_steps))
        assert c == cls[y, x], (x, y, c, cls[y, x])
        logs.append(log)
        lens.append(len(log))
        stats.append([st["attempts"], st["rhs"], st["accepted"], st["crossings"], st["steps"]])
    return dict(hdr=hdr, cls=cls, attempts=att, log_pixels=np.array(pixels, dtype=np.int32),
                log_len=np.array(lens, dtype=np.int32), logs=np.concatenate(logs, axis=0),
                log_stats=np.array(stats, dtype=np.uint32))


def main():
    arrays = make()
    np.savez_compressed(OUT, **arrays)
    print(OUT, {k: v.shape for k, v in arrays.items()}, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
