"""Frame output: the first-frame PNG of the reference's render loop and raw
array dumps of a frame for parity tooling.

``save_first_frame`` follows src/app.rs:219-237: the ready frame's
``0x00RRGGBB`` pixels become RGBA with alpha 255, written as an 8-bit RGBA
PNG, and an existing file is left alone.  The PNG encoder is written here
(stdlib zlib; one IDAT, filter type 0 on every row) so the library needs
no imaging package.
"""
import os
import struct
import zlib

import numpy as np

_PNG_SIGNATURE = b"\x89PNG\r\n\x1a\n"


def ldr_to_rgba(ldr):
    """uint8 [H, W, 4] RGBA (alpha 255) from a [H, W] array of 0x00RRGGBB
    pixels (uint32 or int32, e.g. the frame API's view or post_process's
    output)."""
    px = np.asarray(ldr).astype(np.uint32, copy=False)
    if px.ndim != 2:
        raise ValueError(f"expected a [H, W] pixel array, got shape {px.shape}")
    rgba = np.empty(px.shape + (4,), dtype=np.uint8)
    rgba[..., 0] = (px >> 16) & 0xFF
    rgba[..., 1] = (px >> 8) & 0xFF
    rgba[..., 2] = px & 0xFF
    rgba[..., 3] = 255
    return rgba


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def encode_png(rgba, level=6):
    """PNG bytes of an 8-bit RGBA image [H, W, 4]."""
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError(f"expected [H, W, 4] uint8 RGBA, got shape {img.shape}")
    h, w = img.shape[:2]
    if h == 0 or w == 0 or h > 0x7FFFFFFF or w > 0x7FFFFFFF:
        raise ValueError(f"image size {w}x{h} is not a valid PNG size")
    rows = np.empty((h, 1 + 4 * w), dtype=np.uint8)
    rows[:, 0] = 0  # filter type 0 (none) on every scanline
    rows[:, 1:] = img.reshape(h, 4 * w)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)  # 8-bit, RGBA, deflate, no interlace
    return (_PNG_SIGNATURE + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(rows.tobytes(), level)) +
            _chunk(b"IEND", b""))


def write_png(path, rgba):
    """Write an 8-bit RGBA image [H, W, 4] as a PNG file."""
    data = encode_png(rgba)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(data)


def save_first_frame(path, ldr):
    """The reference's first-frame capture (src/app.rs:219-237): write the
    frame as RGBA PNG unless `path` already exists.  Returns True if a file
    was written."""
    if os.path.exists(path):
        return False
    write_png(path, ldr_to_rgba(ldr))
    return True


def save_frame_arrays(prefix, hdr=None, cls=None, attempts=None, ldr=None):
    """Raw arrays of one frame as .npy files next to each other:
    `<prefix>_hdr.npy` (float32 [H, W, 4]), `<prefix>_class.npy` (uint8),
    `<prefix>_attempts.npy` (uint32), `<prefix>_ldr.npy` (uint32
    0x00RRGGBB).  Accepts numpy arrays or device tensors.  Returns the paths
    written."""
    out = []
    for name, arr, dtype in (("hdr", hdr, np.float32), ("class", cls, np.uint8),
                             ("attempts", attempts, np.uint32), ("ldr", ldr, np.uint32)):
        if arr is None:
            continue
        if hasattr(arr, "detach"):
            arr = arr.detach().cpu().numpy()
        a = np.asarray(arr)
        if a.dtype != dtype:
            a = a.view(dtype) if a.dtype.itemsize == np.dtype(dtype).itemsize else a.astype(dtype)
        p = f"{prefix}_{name}.npy"
        os.makedirs(os.path.dirname(os.path.abspath(p)), exist_ok=True)
        np.save(p, a, allow_pickle=False)
        out.append(p)
    return out
