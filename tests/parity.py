"""Comparison metrics of a GPU frame against the FP64 oracle.

Hit class of a pixel: DISK if any disk emission was composited, else its
terminal reason (HORIZON, ESCAPE, OTHER).  Its sub-class adds the sky-grid
flag and the number of disk contributions (bits 3 and 5-7 of the class
byte).  A boundary pixel is
one whose oracle 3x3 neighbourhood holds more than one class or
contribution count, or an escape ray on a sky grid line in either frame
(the grid test is a hard threshold on atan2/asin).  A pixel whose sub-class differs between
the two frames is a sub-class mismatch; colour is compared only where the
sub-classes agree.
Relative error per channel: |d| / max(|ref|, 1e-6 * max|ref|).
"""
import numpy as np

DISK, HORIZON, ESCAPE, OTHER = 3, 1, 2, 0


def hit_class(cls):
    c = np.asarray(cls).astype(np.int32)
    out = c & 3
    out = np.where((c & 4) != 0, DISK, out)
    return out


def sub_class(cls):
    c = np.asarray(cls).astype(np.int32)
    return hit_class(c) * 16 + (c & 8) + ((c >> 5) & 7)


def boundary_mask(ref_cls):
    hc = sub_class(ref_cls)
    H, W = hc.shape
    pad = np.pad(hc, 1, mode="edge")
    diff = np.zeros_like(hc, dtype=bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            diff |= pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] != hc
    grid = (np.asarray(ref_cls).astype(np.int32) & 8) != 0
    return diff | grid


def compare(gpu_hdr, gpu_cls, ref_hdr, ref_cls):
    """Returns a dict of parity metrics.  gpu_hdr: [H,W,3|4]; ref_hdr: [H,W,3]."""
    g = np.asarray(gpu_hdr, dtype=np.float64)[..., :3]
    r = np.asarray(ref_hdr, dtype=np.float64)[..., :3]
    hg, hr = hit_class(gpu_cls), hit_class(ref_cls)
    bnd = boundary_mask(ref_cls) | ((np.asarray(gpu_cls).astype(np.int32) & 8) != 0)
    mism = hg != hr
    sub_mism = sub_class(gpu_cls) != sub_class(ref_cls)
    same = ~sub_mism
    scale = max(float(np.abs(r).max()), 1e-300)
    den = np.maximum(np.abs(r), 1e-6 * scale)
    rel = np.abs(g - r) / den
    rel_px = rel.max(axis=-1)
    rel_same = rel_px[same]
    n = hg.size
    return {
        "pixels": int(n),
        "class_mismatch": int(mism.sum()),
        "class_mismatch_non_boundary": int((mism & ~bnd).sum()),
        "class_mismatch_frac": float(mism.sum()) / n,
        "subclass_mismatch": int(sub_mism.sum()),
        "subclass_mismatch_non_boundary": int((sub_mism & ~bnd).sum()),
        "boundary_frac": float(bnd.sum()) / n,
        "max_rel": float(rel_same.max()) if rel_same.size else 0.0,
        "rms_rel": float(np.sqrt(np.mean(rel_same ** 2))) if rel_same.size else 0.0,
        "max_rel_non_boundary": float(rel_px[same & ~bnd].max()) if (same & ~bnd).any() else 0.0,
        "p999_rel": float(np.quantile(rel_same, 0.999)) if rel_same.size else 0.0,
        "exact_fp32_frac": float(np.mean(np.all(np.asarray(gpu_hdr)[..., :3] == r.astype(np.float32), axis=-1))),
    }


def ldr_diff(a, b):
    """Largest per-channel byte difference of two 0x00RRGGBB images."""
    a = np.asarray(a).astype(np.int64) & 0xFFFFFF
    b = np.asarray(b).astype(np.int64) & 0xFFFFFF
    d = 0
    hist = {}
    for sh in (16, 8, 0):
        dc = np.abs(((a >> sh) & 255) - ((b >> sh) & 255))
        d = max(d, int(dc.max()))
        for v, c in zip(*np.unique(dc, return_counts=True)):
            hist[int(v)] = hist.get(int(v), 0) + int(c)
    return d, hist
