"""Writes label_tex_crop.npz from render.rs's own texture fixture
(tests/fixtures/label.tex, 256 x 256, 9 mip levels written by its tex.rs).

Kept: a 64 x 64 crop of level 0 at texel (64, 64) and, for levels 1-6, the
blocks that crop alone determines (256 is a power of two, so no odd-size
clamping reaches them): level k's block at (64 >> k, 64 >> k), 64 >> k on a
side, bit for bit as the file holds them. Run from the repository root
where the reference tree is present:

    python3 tests/golden/make_tex_golden.py /root/reference/tests/fixtures/label.tex
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "..", "render.rs_amd"))
import rtex  # noqa: E402


def main(path):
    levels = rtex.read_tex(path, levels=True)
    out = {"level0": levels[0][64:128, 64:128].copy()}
    for k in range(1, 7):
        o, n = 64 >> k, 64 >> k
        out[f"level{k}"] = levels[k][o:o + n, o:o + n].copy()
    dst = os.path.join(os.path.dirname(__file__), "label_tex_crop.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: " + ", ".join(f"{k} {v.shape}" for k, v in out.items()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/fixtures/label.tex")
