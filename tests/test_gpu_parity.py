"""Bit-exact parity of the HIP codec with the CPU oracle, through the C-ABI.

Codestream regions (MSB word, codewords, -1 fill) must be identical to the oracle's,
and decoding must give back the original coefficients exactly.
"""
import numpy as np
import pytest

import synth
from conftest import lut_folder

pytestmark = pytest.mark.gpu


def first_mismatch(a, b, stride):
    d = np.nonzero(a.ravel() != b.ravel())[0]
    if d.size == 0:
        return None
    k = int(d[0])
    return "first mismatch at word %d (codeblock %d, slot %d): gpu %d oracle %d; %d words differ" % (
        k, k // stride, k % stride, a.ravel()[k], b.ravel()[k], d.size)


def check_codec(bpc, O, img, lut_arr, cp3, levels=5, rows=64, nbp=15, stride=0, layout=0, nthreads=0):
    Y, X = img.shape
    op = O.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp, cb_stride_words=stride)
    st, ostream, oncw = O.encode(img, lut_arr, op, nthreads)
    assert st == 0
    p = bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp, cb_stride_words=stride,
                   layout=layout)
    lut = bpc.Lut.from_array(lut_arr)
    try:
        src = img if layout == 0 else synth.to_tiled(img, rows)
        d_in = bpc.DeviceBuffer.from_array(np.ascontiguousarray(src).ravel())
        nw = bpc.stream_words(X, Y, p)
        d_cs = bpc.DeviceBuffer(nw)
        d_ncw = bpc.DeviceBuffer(p.n_codeblocks(X, Y))
        d_out = bpc.DeviceBuffer(X * Y)
        bpc.encode_device(d_in, d_cs, X, Y, lut, p, 1, d_ncw)
        assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
        gstream = d_cs.download()
        msg = first_mismatch(gstream, ostream, p.stride())
        assert msg is None, msg
        assert np.array_equal(d_ncw.download(), oncw)
        bpc.decode_device(d_cs, d_out, X, Y, lut, p)
        assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
        rec = d_out.download()
        rec = rec.reshape(Y, X) if layout == 0 else synth.from_tiled(rec, X, Y, rows)
        assert np.array_equal(rec, img)
        for b in (d_in, d_cs, d_ncw, d_out):
            b.free()
    finally:
        lut.free()
    return ostream, oncw


@pytest.mark.parametrize("cpass3", [0, 1])
def test_sample_image_drop_in(bpc, O, sample_image, cpass3):
    """The reference's own check (main_example.cu:516-532) through the drop-in entry
    point, plus equality of the codestream with the oracle's."""
    lut = bpc.Lut.from_text(lut_folder(cpass3), cpass3)
    try:
        p = bpc.params(cpass3=cpass3, lut_n_bitplanes=lut.nbp)
        bpc.set_config(p)
        buf = sample_image.copy().ravel()
        bpc.run(bpc.CODE, buf, 2048, 2048, lut)
        olut, nbp = O.lut_load_text(lut_folder(cpass3), cpass3)
        st, ostream, _ = O.encode(sample_image, olut, O.params(cpass3=cpass3, lut_n_bitplanes=nbp))
        assert st == 0
        msg = first_mismatch(buf, ostream, 4096)
        assert msg is None, msg
        bpc.run(bpc.DECODE, buf, 2048, 2048, lut)
        assert np.array_equal(buf.reshape(2048, 2048), sample_image)
    finally:
        bpc.set_config(bpc.params())
        lut.free()


def trained(bpc, size, cp3, nbp, seed, levels=5, mag_bits=14, ll_bits=13, detail_exp=9):
    img = synth.synth_image(size, size, levels, seed, ll_bits=ll_bits, detail_exp=detail_exp, mag_bits=mag_bits)
    train = synth.synth_image(size, size, levels, seed + 1, ll_bits=ll_bits, detail_exp=detail_exp, mag_bits=mag_bits)
    lut = bpc.train_lut(train, bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp))
    return img, lut


def test_plain_2048_2pass(bpc, O):
    img, lut = trained(bpc, 2048, 0, 15, seed=100)
    check_codec(bpc, O, img, lut, 0, nbp=15, layout=1)


def test_plain_4096_3pass(bpc, O):
    img, lut = trained(bpc, 4096, 1, 16, seed=200, mag_bits=16)
    check_codec(bpc, O, img, lut, 1, nbp=16, layout=1)


@pytest.mark.parametrize("cp3,nbp,mag", [(1, 16, 16), (0, 20, 19)])
def test_plain_6144_more_waves_than_resident(bpc, O, cp3, nbp, mag):
    """6144^2 needs more waves than fit at once with the bitplanes in registers, so
    the per-image launch keeps them in its scratch slots and grid-strides (16- and
    20-plane kernels); 4096^2 above runs the register form."""
    img, lut = trained(bpc, 6144, cp3, nbp, seed=250 + nbp, mag_bits=mag, ll_bits=min(13, mag),
                       detail_exp=9 if mag <= 16 else mag - 6)
    check_codec(bpc, O, img, lut, cp3, nbp=nbp, layout=1, stride=0 if mag <= 16 else 8192)


@pytest.mark.parametrize("size_y", [2048, 2112])
def test_one_codeblock_per_wave_and_pairs(bpc, O, size_y):
    """The per-image launch codes one codeblock per wave when the image has no more
    codeblocks than the device has SIMDs (1024 on MI355X: 2048x2048), and pairs
    above that (2048x2112 = 1056 codeblocks); both are the oracle's codestream."""
    print("mode threshold", size_y, flush=True)
    img = synth.synth_image(2048, size_y, 5, 230 + size_y, mag_bits=15)
    p = bpc.params(cpass3=1, lut_n_bitplanes=15)
    lut = bpc.train_lut(synth.synth_image(2048, size_y, 5, 231 + size_y, mag_bits=15), p)
    check_codec(bpc, O, img, lut, 1, nbp=15)


@pytest.mark.parametrize("cp3", [0, 1])
def test_3072_straddling_subbands(bpc, O, cp3):
    """3072 = 96 * 32: the LL band is 96x96, so codeblocks straddle subbands and the
    lanes of one codeblock use different LUT groups (BPC_PaCo.cu:1126-1127,1141)."""
    img, lut = trained(bpc, 3072, cp3, 15, seed=300 + cp3)
    check_codec(bpc, O, img, lut, cp3, nbp=15)


@pytest.mark.parametrize("rows,levels,cp3", [(32, 3, 0), (16, 2, 1), (2, 0, 0), (8, 1, 1), (64, 0, 1)])
def test_codeblock_lengths_and_levels(bpc, O, rows, levels, cp3):
    rng = np.random.default_rng(rows * 10 + levels)
    Y, X = 4 * rows * 4, 512
    img = synth.synth_image(X, Y, levels, seed=rows, mag_bits=12) if levels else \
        (np.floor(rng.exponential(40, (Y, X))) * rng.choice([-1, 1], (Y, X))).astype(np.int32)
    p = bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=15)
    lut = bpc.train_lut(img, p)
    check_codec(bpc, O, img, lut, cp3, levels=levels, rows=rows, nbp=15)
    check_codec(bpc, O, img, lut, cp3, levels=levels, rows=rows, nbp=15, layout=1)


@pytest.mark.parametrize("cp3", [0, 1])
def test_odd_codeblock_count_and_zero_blocks(bpc, O, cp3):
    """3 codeblocks (the second wave64 half of the last wave is empty), one all zero
    (MSB word 32, nothing coded), one with a single nonzero coefficient."""
    img = np.zeros((64, 192), np.int32)
    img[:, 128:] = np.random.default_rng(5).integers(-300, 300, (64, 64))
    img[17, 3] = -1
    lut = np.random.default_rng(6).integers(1, 128, 3360 + (3120 if cp3 else 0)).astype(np.int32)
    ostream, oncw = check_codec(bpc, O, img, lut, cp3, levels=0)
    assert ostream[4096] == 32 and oncw[1] == 0


@pytest.mark.parametrize("cp3", [0, 1])
def test_wide_magnitudes(bpc, O, cp3):
    """Magnitudes up to 2^27 - 1 (MSB 27, the largest the word format holds) with a
    28-bitplane LUT: the 28-plane kernel variant.  ~27 bits per sample need more than
    the reference's 4095 codeword slots, so the regions are 16384 words."""
    rng = np.random.default_rng(9)
    img = rng.integers(-(1 << 27) + 1, 1 << 27, (128, 128)).astype(np.int32)
    img[64:, 64:] >>= 20
    p = bpc.params(cpass3=cp3, dwt_levels=1, lut_n_bitplanes=28)
    lut = bpc.train_lut(img, p)
    check_codec(bpc, O, img, lut, cp3, levels=1, nbp=28, stride=16384)


def test_msb_beyond_lut_planes_aliases_like_reference(bpc, O):
    """MSB >= LUT_N_BITPLANES reads the next group's rows, as the reference's pointer
    arithmetic does (BPC_PaCo.cu:228-258); the drop-in entry reruns with wider
    bitplane registers, the device-resident entry reports BPC_ERR_WIDE."""
    rng = np.random.default_rng(10)
    img = rng.integers(-(1 << 17), 1 << 17, (64, 128)).astype(np.int32)  # MSB 17 in both codeblocks
    img[:, :64] >>= 12  # codeblock 0 (LL) MSB 5
    p = bpc.params(cpass3=0, dwt_levels=1, lut_n_bitplanes=15, cb_stride_words=16384)
    lut_arr = rng.integers(1, 128, bpc.lut_len(p)).astype(np.int32)
    op = O.params(cpass3=0, dwt_levels=1, lut_n_bitplanes=15, cb_stride_words=16384)
    # codeblock 1 is level-0 HL (group 0): rows up to 17 stay inside the LUT array
    st, ostream, _ = O.encode(img, lut_arr, op)
    assert st == 0
    lut = bpc.Lut.from_array(lut_arr)
    try:
        buf = np.zeros(2 * 16384, np.int32)
        buf[:img.size] = img.ravel()
        bpc.run(bpc.CODE, buf, 128, 64, lut, p)
        assert np.array_equal(buf, ostream)
        bpc.run(bpc.DECODE, buf, 128, 64, lut, p)
        assert np.array_equal(buf[:img.size].reshape(64, 128), img)
        d_in = bpc.DeviceBuffer.from_array(img.ravel())
        d_cs = bpc.DeviceBuffer(2 * 16384)
        bpc.encode_device(d_in, d_cs, 128, 64, lut, p)
        assert bpc.status_take() == bpc.ERR_WIDE
    finally:
        lut.free()


def test_error_paths(bpc, O):
    rng = np.random.default_rng(11)
    p = bpc.params(cpass3=0, dwt_levels=0)
    lut = bpc.Lut.from_array(np.full(bpc.lut_len(p), 64, np.int32))
    try:
        # |v| >= 2^28 collides with the flag bits
        img = rng.integers(-50, 50, (64, 64)).astype(np.int32)
        img[3, 3] = 1 << 28
        with pytest.raises(bpc.BPCError) as e:
            bpc.run(bpc.CODE, img.copy().ravel(), 64, 64, lut, p)
        assert e.value.code == bpc.ERR_RANGE
        # overflow: a 64-word region cannot hold a noisy codeblock's codewords
        img = rng.integers(-5000, 5000, (64, 64)).astype(np.int32)
        small = bpc.params(cpass3=0, dwt_levels=0, cb_stride_words=64)
        with pytest.raises(bpc.BPCError) as e:
            bpc.run(bpc.CODE, img.copy().ravel(), 64, 64, lut, small)
        assert e.value.code == bpc.ERR_OVERFLOW
        st, _, _ = O.encode(img, np.full(bpc.lut_len(p), 64, np.int32),
                            O.params(cpass3=0, dwt_levels=0, cb_stride_words=64))
        assert st == O.ERR_OVERFLOW
        # malformed MSB word for the decoder
        cs = np.full(4096, -1, np.int32)
        cs[0] = 29
        with pytest.raises(bpc.BPCError) as e:
            bpc.run(bpc.DECODE, cs, 64, 64, lut, p)
        assert e.value.code == bpc.ERR_MALFORMED
        # all-zero sentinel decodes to zeros
        cs[0] = 32
        bpc.run(bpc.DECODE, cs, 64, 64, lut, p)
        assert (cs == 0).all()
    finally:
        lut.free()


def test_unfilled_regions_keep_codewords(bpc, O):
    """fill_unused=0 writes only word 0 and the used slots of each region."""
    img, lut_arr = trained(bpc, 1024, 1, 15, seed=400)
    p = bpc.params(cpass3=1, lut_n_bitplanes=15)
    st, ostream, oncw = O.encode(img, lut_arr, O.params(cpass3=1, lut_n_bitplanes=15))
    lut = bpc.Lut.from_array(lut_arr)
    try:
        d_in = bpc.DeviceBuffer.from_array(img.ravel())
        d_cs = bpc.DeviceBuffer(1024 * 1024)
        d_cs.fill_bytes(0x5A)
        bpc.encode_device(d_in, d_cs, 1024, 1024, lut, p, fill_unused=0)
        assert bpc.status_take() == 0
        g = d_cs.download().reshape(-1, 4096)
        o = ostream.reshape(-1, 4096)
        for cb in range(g.shape[0]):
            n = oncw[cb]
            assert np.array_equal(g[cb, :n + 1], o[cb, :n + 1])
            assert (g[cb, n + 1:] == 0x5A5A5A5A).all()
    finally:
        lut.free()
