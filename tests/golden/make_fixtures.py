"""Writes the committed fixtures under tests/golden/ from the reference's data files.

Run once in a container that has the reference checkout (it reads only data:
sample_image/sample_image.zip and the LUT text files; no reference code runs):

    python tests/golden/make_fixtures.py /root/reference

Outputs
  sample_image.i16.xz      2048x2048 raster (row-major, little-endian int16, lzma) of the
                           reference's sample image, with codeblock k.txt placed where
                           load_Data (main_example.cu:71-179) places it
  sample_headers.json      per codeblock file k: header (level, subband, MSBP) and the
                           (x, y) position load_Data gives it
  luts/{2,3}_coding_pass/  the reference's probability LUT text files, copied verbatim
"""
import io
import json
import lzma
import os
import shutil
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_data_positions(size_x, size_y, levels, cbw=64, cbh=64):
    """Codeblock placement order of main_example.cu:71-179: the LL band first, then
    for each level from coarsest to finest the HL (right), LH (below), HH bands,
    each scanned in row-major codeblock order."""
    pos = []
    sx, sy = size_x >> levels, size_y >> levels
    for j in range(0, sy, cbh):
        for i in range(0, sx, cbw):
            pos.append((i, j))
    for d in range(levels, 0, -1):
        sx, sy = size_x >> d, size_y >> d
        for (x0, y0) in ((sx, 0), (0, sy), (sx, sy)):
            for j in range(0, sy, cbh):
                for i in range(0, sx, cbw):
                    pos.append((x0 + i, y0 + j))
    return pos


def main(ref_root):
    size = 2048
    levels = 5
    zpath = os.path.join(ref_root, "sample_image", "sample_image.zip")
    img = np.zeros((size, size), dtype=np.int16)
    headers = []
    pos = load_data_positions(size, size, levels)
    with zipfile.ZipFile(zpath) as z:
        for k, (x, y) in enumerate(pos):
            text = z.read("%d.txt" % k).decode()
            first, rest = text.split("\n", 1)
            h = first.split()  # "level subband MSBP: k"
            vals = np.array(rest.split(), dtype=np.int32)
            assert vals.size == 64 * 64, (k, vals.size)
            assert np.abs(vals).max() < 2 ** 15
            img[y:y + 64, x:x + 64] = vals.reshape(64, 64).astype(np.int16)
            headers.append({"k": k, "level": int(h[0]), "subband": int(h[1]), "msbp": int(h[3]),
                            "x": x, "y": y})
    with open(os.path.join(HERE, "sample_image.i16.xz"), "wb") as f:
        f.write(lzma.compress(img.astype("<i2").tobytes(), preset=9 | lzma.PRESET_EXTREME))
    with open(os.path.join(HERE, "sample_headers.json"), "w") as f:
        json.dump({"size_x": size, "size_y": size, "dwt_levels": levels, "codeblocks": headers}, f,
                  separators=(",", ":"))
    for sub in ("2_coding_pass", "3_coding_pass"):
        src = os.path.join(ref_root, "LUTs", sub)
        dst = os.path.join(HERE, "luts", sub)
        os.makedirs(dst, exist_ok=True)
        for name in sorted(os.listdir(src)):
            shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
