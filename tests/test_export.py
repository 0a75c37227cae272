"""Frame output (src/app.rs:219-237): 0x00RRGGBB -> RGBA with alpha 255,
8-bit RGBA PNG, no overwrite of an existing first frame; raw array dumps."""
import struct
import zlib

import numpy as np
import pytest

from kerrtrace_amd import export as E


def _decode_png(data):
    """Minimal PNG reader for the files encode_png writes: chunk CRCs,
    IHDR fields, inflate, filter type 0 rows."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        chunks.append((kind, body))
        pos += 12 + n
    assert [k for k, _ in chunks][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for k, b in chunks if k == b"IDAT"))
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def _ldr(h, w, seed=0):
    return np.random.default_rng(seed).integers(0, 1 << 24, (h, w), dtype=np.uint32)


def test_ldr_to_rgba_channels_and_alpha():
    px = np.array([[0x00FF8001, 0x00000000], [0x00123456, 0x00FFFFFF]], dtype=np.uint32)
    rgba = E.ldr_to_rgba(px)
    assert rgba.dtype == np.uint8 and rgba.shape == (2, 2, 4)
    assert rgba[0, 0].tolist() == [0xFF, 0x80, 0x01, 255]
    assert rgba[1, 0].tolist() == [0x12, 0x34, 0x56, 255]
    assert np.all(rgba[..., 3] == 255)
    # int32 views of the same pixels (torch has no uint32) give the same image
    assert np.array_equal(E.ldr_to_rgba(px.view(np.int32)), rgba)


def test_png_round_trip():
    ldr = _ldr(37, 53)
    rgba = E.ldr_to_rgba(ldr)
    back = _decode_png(E.encode_png(rgba))
    assert np.array_equal(back, rgba)


def test_png_readable_by_pil(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    ldr = _ldr(21, 40, seed=3)
    p = tmp_path / "f.png"
    E.write_png(str(p), E.ldr_to_rgba(ldr))
    im = Image.open(str(p))
    assert im.mode == "RGBA" and im.size == (40, 21)
    assert np.array_equal(np.asarray(im), E.ldr_to_rgba(ldr))


def test_save_first_frame_does_not_overwrite(tmp_path):
    p = tmp_path / "picture" / "first_frame.png"
    assert E.save_first_frame(str(p), _ldr(8, 8, seed=1))
    first = p.read_bytes()
    assert not E.save_first_frame(str(p), _ldr(8, 8, seed=2))
    assert p.read_bytes() == first


def test_encode_rejects_bad_shapes():
    with pytest.raises(ValueError):
        E.encode_png(np.zeros((4, 4, 3), dtype=np.uint8))
    with pytest.raises(ValueError):
        E.encode_png(np.zeros((0, 4, 4), dtype=np.uint8))
    with pytest.raises(ValueError):
        E.ldr_to_rgba(np.zeros((2, 2, 2), dtype=np.uint32))


def test_frame_arrays(tmp_path):
    hdr = np.random.default_rng(0).random((5, 6, 4)).astype(np.float32)
    cls = np.arange(30, dtype=np.uint8).reshape(5, 6)
    ldr = _ldr(5, 6).view(np.int32)
    paths = E.save_frame_arrays(str(tmp_path / "frame"), hdr=hdr, cls=cls, ldr=ldr)
    assert [p.rsplit("_", 1)[1] for p in paths] == ["hdr.npy", "class.npy", "ldr.npy"]
    assert np.array_equal(np.load(paths[0]), hdr)
    assert np.array_equal(np.load(paths[1]), cls)
    assert np.array_equal(np.load(paths[2]), ldr.view(np.uint32))
