"""Writes diagram_first_frame.json, the summary of the reference's only
artefact, picture/diagram.png (2100x900 RGBA8, the first frame the app
saves at src/app.rs:207-237), that tests/diagram_check.py compares
rendered frames with: the 210x90 grid of 10x10 block means of the Rec. 709
luma of its 8-bit sRGB pixels (rounded, base64 of uint8, row-major) and
the 16-bin histogram of its per-pixel luma.  Reads the image from the
reference tree, so it runs only where that tree exists.
"""
import base64
import json
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.diagram_check import FIXTURE, GRID, summary  # noqa: E402

SOURCE = "/root/reference/picture/diagram.png"


def main():
    im = Image.open(SOURCE)
    assert im.size == (2100, 900), im.size
    rgba = np.asarray(im.convert("RGBA"))
    assert np.all(rgba[..., 3] == 255)
    blocks, hist = summary(rgba[..., :3])
    assert blocks.shape == GRID
    out = {"source": "picture/diagram.png", "size": [2100, 900], "grid": list(GRID),
           "block_luma_u8": base64.b64encode(blocks.tobytes()).decode(),
           "hist16": [round(float(v), 8) for v in hist]}
    with open(FIXTURE, "w") as fh:
        json.dump(out, fh)
    print(FIXTURE, os.path.getsize(FIXTURE), "bytes")


if __name__ == "__main__":
    main()
