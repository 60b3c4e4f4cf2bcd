"""The oracle against the reference's own data: the sample image's codeblock headers
(level, subband, MSBP of all 1024 files) and the reference's self-check
(main_example.cu:518-532): encode, decode, compare with the original."""
import numpy as np
import pytest

from conftest import lut_folder


def test_headers_pin_msb_and_subband(O, sample_image, sample_headers):
    X = sample_headers["size_x"]
    L = sample_headers["dwt_levels"]
    nx = X // 64
    for h in sample_headers["codeblocks"]:
        x, y = h["x"], h["y"]
        cb = (y // 64) * nx + x // 64
        assert O.codeblock_msb(sample_image, cb) == h["msbp"], h
        # every lane of the codeblock classifies to the header's band (BPC_PaCo.cu:1126-1127,1141)
        for lane in range(32):
            assert O.find_subband(x + 2 * lane, y, X, X, L) == (h["level"], h["subband"]), h


@pytest.mark.parametrize("cpass3", [0, 1])
def test_sample_image_round_trip(O, sample_image, cpass3):
    lut, nbp = O.lut_load_text(lut_folder(cpass3), cpass3)
    p = O.params(cpass3=cpass3, lut_n_bitplanes=nbp)
    st, stream, ncw = O.encode(sample_image, lut, p)
    assert st == 0
    st, rec = O.decode(stream, 2048, 2048, lut, p)
    assert st == 0
    assert np.array_equal(rec, sample_image)
    # layout invariants: region word 0 = MSB, then ncw codewords < 2^16, then -1
    regions = stream.reshape(-1, 4096)
    assert (ncw >= 1).all() and (ncw <= 4095).all()
    for cb in range(0, regions.shape[0], 37):
        r = regions[cb]
        assert 0 <= r[0] <= 27
        n = ncw[cb]
        assert (r[1:1 + n] >= 0).all() and (r[1:1 + n] <= 0xFFFF).all()
        assert (r[1 + n:] == -1).all()


def test_sample_image_compresses(O, sample_image):
    lut, nbp = O.lut_load_text(lut_folder(0), 0)
    st, stream, ncw = O.encode(sample_image, lut, O.params(lut_n_bitplanes=nbp))
    bits_per_sample = ncw.sum() * 16 / sample_image.size
    # lossless coding of 13-bit-range DWT data with trained LUTs: well under raw size
    assert 3.0 < bits_per_sample < 8.0
