"""Large image through the batched persistent path: 16384x16384 (65536 codeblocks,
1 GiB of int32), 20 magnitude bitplanes (LL uniform on [0, 2^19), detail scale
2^(14-d)), 3 passes, 8192-word regions.  Bit-exact against the oracle on a seeded
1% sample of codeblocks, plus equality of the full-image round trip."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu

SIZE = 16384
NBP = 20
STRIDE = 8192


def test_large_16384_sampled_parity_and_round_trip(bpc, O):
    img = synth.synth_image(SIZE, SIZE, 5, 77, ll_bits=19, detail_exp=14, mag_bits=NBP)
    train = synth.synth_image(SIZE // 4, SIZE // 4, 5, 78, ll_bits=17, detail_exp=12, mag_bits=NBP)
    lut_arr = bpc.train_lut(train, bpc.params(cpass3=1, lut_n_bitplanes=NBP))
    del train
    p = bpc.params(cpass3=1, lut_n_bitplanes=NBP, cb_stride_words=STRIDE, layout=bpc.LAYOUT_TILED)
    ncb = p.n_codeblocks(SIZE, SIZE)
    assert ncb == 65536
    lut = bpc.Lut.from_array(lut_arr)
    d_in = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
    d_cs = bpc.DeviceBuffer(bpc.stream_words(SIZE, SIZE, p))
    d_out = bpc.DeviceBuffer(SIZE * SIZE)
    d_ncw = bpc.DeviceBuffer(ncb)
    batch = bpc.Batch(1, ncb)
    try:
        batch.run(bpc.CODE, [bpc.make_image(d_in, d_cs, lut, SIZE, SIZE, p, d_ncw)])
        batch.run(bpc.DECODE, [bpc.make_image(d_cs, d_out, lut, SIZE, SIZE, p)])
        assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
        rec = synth.from_tiled(d_out.download(), SIZE, SIZE)
        assert np.array_equal(rec, img)
        del rec
        ncw = d_ncw.download()
        assert ncw.max() < STRIDE - 1 and ncw.max() > 4095  # needs more than the reference's region
        cs = d_cs.download()
        rng = np.random.default_rng(2024)
        sample = np.sort(rng.choice(ncb, size=ncb // 100, replace=False)).astype(np.int32)
        op = O.params(cpass3=1, lut_n_bitplanes=NBP, cb_stride_words=STRIDE)
        st, ostream, oncw = O.encode_list(img, sample, lut_arr, op)
        assert st == 0
        regions = cs.reshape(ncb, STRIDE)
        oregions = ostream.reshape(ncb, STRIDE)
        for cb in sample:
            assert np.array_equal(regions[cb], oregions[cb]), cb
        assert np.array_equal(ncw[sample], oncw[sample])
        st, orec = O.decode_list(cs, SIZE, SIZE, sample, lut_arr, op)
        assert st == 0
        nx = SIZE // 64
        for cb in sample[:64]:
            ty, tx = divmod(int(cb), nx)
            blk = np.s_[ty * 64:(ty + 1) * 64, tx * 64:(tx + 1) * 64]
            assert np.array_equal(orec[blk], img[blk]), cb
    finally:
        batch.destroy()
        for b in (d_in, d_cs, d_out, d_ncw):
            b.free()
        lut.free()
