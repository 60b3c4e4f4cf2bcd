"""Device LUT trainer: per-entry symbol counts identical to the oracle's, and the
resulting table identical to the host trainer's, through the C-ABI."""
import numpy as np
import pytest

import synth
from test_oracle_counts import counts_image, lut_from_counts

pytestmark = pytest.mark.gpu


def gpu_counts(bpc, img, p):
    src = img if p.layout == 0 else synth.to_tiled(img, p.cblock_length).reshape(img.shape)
    return bpc.train_lut_gpu(src, p, counts=True)


@pytest.mark.parametrize("cp3,levels,rows,nbp,layout", [
    (0, 2, 64, 8, 0), (1, 2, 64, 8, 0), (1, 3, 32, 14, 0), (0, 0, 64, 15, 0), (1, 5, 64, 15, 1), (0, 5, 16, 15, 1)])
def test_counts_match_oracle(bpc, O, cp3, levels, rows, nbp, layout):
    img = counts_image(levels, rows, seed=70 + cp3 + levels)
    print("counts", cp3, levels, rows, nbp, layout, flush=True)
    op = O.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp)
    st, on0, on = O.symbol_counts(img, op)
    assert st == 0
    p = bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp, layout=layout)
    lut, n0, n = gpu_counts(bpc, img, p)
    assert np.array_equal(n, on), "symbol counts differ at %d entries" % int((n != on).sum())
    assert np.array_equal(n0, on0)
    assert np.array_equal(lut, lut_from_counts(on0, on))


def test_odd_codeblock_count(bpc, O):
    """A grid with an odd number of codeblocks (the last wave codes one)."""
    img = synth.synth_image(192, 192, 2, 81, mag_bits=13)  # 3 x 3 codeblocks
    op = O.params(cpass3=1, dwt_levels=2, lut_n_bitplanes=14)
    st, on0, on = O.symbol_counts(img, op)
    assert st == 0
    lut, n0, n = gpu_counts(bpc, img, bpc.params(cpass3=1, dwt_levels=2, lut_n_bitplanes=14))
    assert np.array_equal(n, on) and np.array_equal(n0, on0)


@pytest.mark.parametrize("size,cp3,nbp,mag", [(2048, 0, 15, 14), (4096, 1, 16, 16), (2048, 1, 20, 20)])
def test_lut_matches_host_trainer(bpc, O, size, cp3, nbp, mag):
    """Whole-image training: the device table equals the host trainer's, and coding
    with the device-resident table gives the oracle's codestream."""
    print("train", size, cp3, nbp, flush=True)
    dist = dict(mag_bits=mag, ll_bits=19 if mag > 16 else 13, detail_exp=14 if mag > 16 else 9)
    train = synth.synth_image(size, size, 5, 90 + cp3, **dist)
    stride = 16384 if mag > 16 else 0  # 20-bitplane codeblocks need more than 4096 words
    p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp, cb_stride_words=stride)
    host = bpc.train_lut(train, p)
    d_train = bpc.DeviceBuffer.from_array(train)
    lut = bpc.Lut.train_device(d_train, size, size, p)
    try:
        dev = bpc.download_ptr(lut.ptr.value, lut.len)
        assert np.array_equal(dev, host)
        img = synth.synth_image(512, 512, 5, 91 + cp3, **dist)
        op = O.params(cpass3=cp3, lut_n_bitplanes=nbp, cb_stride_words=stride)
        st, ostream, _ = O.encode(img, host, op)
        assert st == 0
        d_img = bpc.DeviceBuffer.from_array(img)
        d_cs = bpc.DeviceBuffer(bpc.stream_words(512, 512, p))
        bpc.encode_device(d_img, d_cs, 512, 512, lut, p)
        assert bpc.status_take() == 0
        assert np.array_equal(d_cs.download(), ostream)
        for b in (d_img, d_cs):
            b.free()
    finally:
        lut.free()
        d_train.free()
