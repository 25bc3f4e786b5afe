"""Writes tests/golden/planar_data.npz from the reference's example input
NADIA/examples/image_files/planar_data.tif (1024 x 1024 uint16 diffraction
pattern, the input of NADIA/examples/PlanarCDI_example.c).  The pixels are
stored unchanged (uint16, compressed); readers convert to the x = column,
y = row orientation NADIA's TIFF reader uses (NADIA/src/io_tiff.c:189-197).
Run once in a checkout where the reference is present:

    python tests/golden/make_planar_data.py /root/reference
"""
import os
import sys

import numpy as np
from PIL import Image


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "NADIA", "examples", "image_files", "planar_data.tif")
    img = np.array(Image.open(src))
    assert img.dtype == np.uint16 and img.shape == (1024, 1024)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planar_data.npz")
    np.savez_compressed(out, rows_cols=img)
    print("wrote", out, img.shape, img.dtype, int(img.max()))


if __name__ == "__main__":
    main()
