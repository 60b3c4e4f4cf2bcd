"""int16 coefficient path: the same codestream as from int32 coefficients of equal
value (and as the oracle's), exact int16 reconstruction, saturation reported."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,cp3,layout,rows", [(2048, 0, 0, 64), (2048, 1, 1, 64), (1024, 1, 0, 32)])
def test_int16_matches_int32_and_oracle(bpc, O, size, cp3, layout, rows):
    print("int16", size, cp3, layout, rows, flush=True)
    img = synth.synth_image(size, size, 5, 200 + cp3, mag_bits=15)
    # int16 extremes, planted in detail codeblocks (the magnitude 2^15 only int16's minimum holds)
    h = size // 2
    img[h, h], img[h + 5, h + 7], img[size - 70, 130], img[size - 1, size - 1] = -32768, 32767, -32767, -1
    p = bpc.params(cpass3=cp3, cblock_length=rows, lut_n_bitplanes=16, layout=layout)
    lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, 201 + cp3, mag_bits=15),
                            bpc.params(cpass3=cp3, cblock_length=rows, lut_n_bitplanes=16))
    st, ostream, oncw = O.encode(img, lut_arr, O.params(cpass3=cp3, cblock_length=rows, lut_n_bitplanes=16))
    assert st == 0
    src = img if layout == 0 else synth.to_tiled(img, rows).reshape(size, size)
    lut = bpc.Lut.from_array(lut_arr)
    ncb = p.n_codeblocks(size, size)
    d16 = bpc.DeviceBuffer.from_array(src.astype(np.int16))
    d32 = bpc.DeviceBuffer.from_array(src.astype(np.int32))
    cs16 = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    cs32 = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    ncw = bpc.DeviceBuffer(ncb)
    out16 = bpc.DeviceBuffer(size * size, np.int16)
    try:
        bpc.encode_device_i16(d16, cs16, size, size, lut, p, 1, ncw)
        bpc.encode_device(d32, cs32, size, size, lut, p)
        assert bpc.status_take() == 0
        g16 = cs16.download()
        assert np.array_equal(g16, cs32.download())
        assert np.array_equal(g16, ostream)
        assert np.array_equal(ncw.download(), oncw)
        out16.fill_bytes(0x5A)
        bpc.decode_device_i16(cs16, out16, size, size, lut, p)
        assert bpc.status_take() == 0
        rec = out16.download().reshape(size, size)
        assert np.array_equal(rec, src.astype(np.int16))
    finally:
        for b in (d16, d32, cs16, cs32, ncw, out16):
            b.free()
        lut.free()


def test_int16_decode_saturates(bpc, O):
    """A codestream of magnitudes past the int16 range decodes saturated, with an error."""
    size = 256
    img = synth.synth_image(size, size, 3, 210, ll_bits=17, mag_bits=17)
    big = np.abs(img.astype(np.int64)) > 32767
    assert big.any()
    p = bpc.params(cpass3=1, dwt_levels=3, lut_n_bitplanes=18)
    lut_arr = bpc.train_lut(img, p)
    st, ostream, _ = O.encode(img, lut_arr, O.params(cpass3=1, dwt_levels=3, lut_n_bitplanes=18))
    assert st == 0
    lut = bpc.Lut.from_array(lut_arr)
    cs = bpc.DeviceBuffer.from_array(ostream)
    out = bpc.DeviceBuffer(size * size, np.int16)
    try:
        bpc.decode_device_i16(cs, out, size, size, lut, p)
        assert bpc.status_take() == bpc.ERR_RANGE
        assert np.array_equal(out.download().reshape(size, size), np.clip(img, -32768, 32767).astype(np.int16))
    finally:
        for b in (cs, out):
            b.free()
        lut.free()
