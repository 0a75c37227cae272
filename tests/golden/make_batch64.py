"""Writes kerrtrace_amd/presets/batch64_seed7.json: 64 square renders drawn once with
numpy.random.default_rng(7).  Per render, in this order:
    spin        ~ U[0, 0.998]
    inclination ~ U[5, 89] deg
    side        = round(exp(U[ln 128, ln 1024])) px  (W = H = side)
    r_obs       ~ U[30, 200] M
The JSON is committed, so loading the batch needs no generator.
Run: python tests/golden/make_batch64.py
"""
import json
import math
import os

import numpy as np

N = 64
SEED = 7


def draw():
    rng = np.random.default_rng(SEED)
    out = []
    for _ in range(N):
        spin = float(rng.uniform(0.0, 0.998))
        incl = float(rng.uniform(5.0, 89.0))
        side = int(round(math.exp(float(rng.uniform(math.log(128.0), math.log(1024.0))))))
        r_obs = float(rng.uniform(30.0, 200.0))
        out.append({"spin": spin, "inclination_deg": incl, "side": side, "r_obs": r_obs})
    return out


def main():
    renders = draw()
    doc = {
        "seed": SEED,
        "generator": "numpy.random.default_rng(7); per render: spin U[0,0.998], inclination U[5,89] deg, "
                     "side round(exp U[ln128, ln1024]), r_obs U[30,200] M",
        "other_knobs": "FOV 20 deg, tol 1e-9, disk ISCO..30 M, FBM 6 octaves seed 1, bloom sigma 12 px",
        "total_rays": sum(r["side"] ** 2 for r in renders),
        "renders": renders,
    }
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    path = os.path.join(root, "kerrtrace_amd", "presets", "batch64_seed7.json")
    with open(path, "w", encoding="utf-8") as fh:
        json.dump(doc, fh, indent=1)
        fh.write("\n")
    print(path, doc["total_rays"])


if __name__ == "__main__":
    main()
