"""Writes tests/golden/noise_ldr_lll1_r.bin: the R channel of the reference's
four blue-noise textures data/noise/LDR_LLL1_{0..3}.png (256x256 RGBA8),
stacked as uint8[4][256][256] with image row 0 first.

These images are the reference's own input data for tilde_h0_k
(main.cpp:1085-1088, tilde_h0_k_cs.glsl:51-68); the fixture lets the
reference-noise mode be tested without the reference tree.  Run once here;
the .bin is committed.
"""
import os

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/data/noise"


def main():
    planes = []
    for i in range(4):
        im = np.asarray(Image.open(os.path.join(SRC, "LDR_LLL1_%d.png" % i)).convert("RGBA"))
        assert im.shape == (256, 256, 4), im.shape
        assert (im[..., 0] == im[..., 1]).all() and (im[..., 3] == 255).all()
        planes.append(im[..., 0].copy())
    arr = np.stack(planes).astype(np.uint8)
    arr.tofile(os.path.join(HERE, "noise_ldr_lll1_r.bin"))
    print("wrote", arr.shape, "histogram min/max count:",
          [int(np.bincount(p.ravel(), minlength=256).min()) for p in arr],
          [int(np.bincount(p.ravel(), minlength=256).max()) for p in arr])


if __name__ == "__main__":
    main()
