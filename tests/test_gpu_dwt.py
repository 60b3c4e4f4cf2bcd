"""5/3 wavelet transform on the device: bit-exact against the oracle in both
directions, and the whole image pipeline (transform, encode, decode, inverse)."""
import numpy as np
import pytest

import dwt53

pytestmark = pytest.mark.gpu


def natural_image(size_x, size_y, seed, bits=12):
    """Smooth structure plus texture and noise, unsigned `bits`-bit samples."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:size_y, 0:size_x].astype(np.float64)
    v = (0.5 + 0.25 * np.sin(x / 37.0) * np.cos(y / 23.0) + 0.15 * np.sin((x + 2 * y) / 7.0)
         + 0.05 * rng.standard_normal((size_y, size_x)))
    return np.clip(v * (1 << bits), 0, (1 << bits) - 1).astype(np.int32)


@pytest.mark.parametrize("size_x,size_y,levels", [
    (2048, 2048, 5), (3072, 3072, 5), (640, 384, 3), (96, 32, 4), (2, 2, 1), (130, 66, 1), (256, 512, 0),
    (4096, 4096, 5)])
def test_transform_matches_oracle(bpc, size_x, size_y, levels):
    print("dwt", size_x, size_y, levels, flush=True)
    rng = np.random.default_rng(size_x + 7 * size_y + levels)
    img = rng.integers(-(1 << 20), 1 << 20, (size_y, size_x)).astype(np.int32)
    want = dwt53.forward(img, levels)
    got = bpc.dwt53(img, levels)
    assert np.array_equal(got, want), "forward differs at %d samples" % int((got != want).sum())
    back = bpc.dwt53(want, levels, inverse=True)
    assert np.array_equal(back, img)


@pytest.mark.parametrize("cp3", [0, 1])
def test_image_pipeline(bpc, O, cp3):
    """image -> forward transform -> encode -> decode -> inverse == image, every
    stage on the device, codestream equal to the oracle's on the oracle's transform."""
    size, levels, nbp = 1024, 5, 16
    img = natural_image(size, size, 300 + cp3)
    train = natural_image(size, size, 301 + cp3)
    p = bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp)
    lut_arr = bpc.train_lut(dwt53.forward(train, levels), p)
    lut = bpc.Lut.from_array(lut_arr)
    d_img = bpc.DeviceBuffer.from_array(img)
    d_coef = bpc.DeviceBuffer(size * size)
    d_cs = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    d_dec = bpc.DeviceBuffer(size * size)
    d_rec = bpc.DeviceBuffer(size * size)
    try:
        bpc.dwt53_forward_device(d_img, d_coef, size, size, levels)
        bpc.encode_device(d_coef, d_cs, size, size, lut, p)
        bpc.decode_device(d_cs, d_dec, size, size, lut, p)
        bpc.dwt53_inverse_device(d_dec, d_rec, size, size, levels)
        assert bpc.status_take() == 0
        assert np.array_equal(d_rec.download().reshape(size, size), img)
        st, ostream, _ = O.encode(dwt53.forward(img, levels), lut_arr, O.params(cpass3=cp3, dwt_levels=levels,
                                                                                 lut_n_bitplanes=nbp))
        assert st == 0
        assert np.array_equal(d_cs.download(), ostream)
    finally:
        for b in (d_img, d_coef, d_cs, d_dec, d_rec):
            b.free()
        lut.free()
