"""Writes tests/golden/sample_codestream.json: the SHA-256 of the oracle's
codestream (int32 little-endian, reference region layout with the -1 fill) for the
reference's sample image with its shipped probability tables, both pass modes,
and the total codeword count.  These digests pin the oracle's output for the
reference's own example (main_example.cu:452-548), so that a later change of the
oracle shows, and the compiled C caller's output is compared against them.

    python tests/golden/make_digests.py
"""
import hashlib
import json
import lzma
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def sample_image():
    with open(os.path.join(HERE, "sample_image.i16.xz"), "rb") as f:
        raw = lzma.decompress(f.read())
    return np.frombuffer(raw, dtype="<i2").reshape(2048, 2048).astype(np.int32)


def digests():
    img = sample_image()
    out = {}
    for cp3 in (0, 1):
        name = "%d_coding_pass" % (3 if cp3 else 2)
        lut, nbp = O.lut_load_text(os.path.join(HERE, "luts", name) + "/", cp3)
        st, cs, ncw = O.encode(img, lut, O.params(cpass3=cp3, lut_n_bitplanes=nbp))
        assert st == 0
        out[name] = {"sha256": hashlib.sha256(cs.astype("<i4").tobytes()).hexdigest(),
                     "codewords": int(ncw.sum()), "lut_n_bitplanes": nbp}
    return out


if __name__ == "__main__":
    d = digests()
    with open(os.path.join(HERE, "sample_codestream.json"), "w") as f:
        json.dump(d, f, indent=1)
    print(d)
