"""Compact codestream container on the device: index / pack / unpack against the
compact form of the oracle's codestream, exact reference-layout round trip, and
per-codeblock files written from the device form."""
import numpy as np
import pytest

import synth
from test_container_host import compact_of

pytestmark = pytest.mark.gpu


def encoded(bpc, O, size, cp3, nbp, seed, rows=64, stride=0, mag_bits=14, zero_cb=True, ll_bits=13):
    img = synth.synth_image(size, size, 5, seed, ll_bits=ll_bits, mag_bits=mag_bits)
    if zero_cb:
        img[0:rows, 64:128] = 0
    p = bpc.params(cpass3=cp3, cblock_length=rows, lut_n_bitplanes=nbp, cb_stride_words=stride)
    lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, seed + 1, ll_bits=ll_bits, mag_bits=mag_bits), p)
    st, ostream, oncw = O.encode(img, lut_arr, O.params(cpass3=cp3, cblock_length=rows, lut_n_bitplanes=nbp,
                                                        cb_stride_words=stride))
    assert st == 0
    return img, p, lut_arr, ostream, oncw


@pytest.mark.parametrize("size,cp3,nbp,rows,stride,derive", [
    (2048, 0, 15, 64, 0, False), (2048, 0, 15, 64, 0, True), (1024, 1, 16, 32, 0, True),
    (1024, 1, 15, 64, 6000, True), (1024, 0, 15, 64, 6002, True), (4096, 1, 16, 64, 0, False)])
def test_pack_unpack_matches_oracle(bpc, O, size, cp3, nbp, rows, stride, derive):
    # regions that are multiples of 4 words move as 16-byte groups; 6002 takes the word path
    print("container", size, cp3, nbp, rows, stride, derive, flush=True)
    img, p, lut_arr, ostream, oncw = encoded(bpc, O, size, cp3, nbp, 140 + size // 1024 + cp3, rows, stride,
                                             mag_bits=16 if nbp >= 16 else 14)
    ncb = p.n_codeblocks(size, size)
    sw = p.stride()
    lut = bpc.Lut.from_array(lut_arr)
    d_img = bpc.DeviceBuffer.from_array(img)
    d_cs = bpc.DeviceBuffer(ncb * sw)
    d_ncw = bpc.DeviceBuffer(ncb)
    d_back = bpc.DeviceBuffer(ncb * sw)
    d_rec = bpc.DeviceBuffer(size * size)
    try:
        bpc.encode_device(d_img, d_cs, size, size, lut, p, 1, d_ncw)
        assert bpc.status_take() == 0
        c = bpc.compact_pack(d_cs, ncb, sw, d_ncw=None if derive else d_ncw)
        assert bpc.status_take() == 0
        msb, offsets, payload = c.download()
        emsb, eoff, epay = compact_of(ostream, oncw, sw)
        assert np.array_equal(msb, emsb)
        assert np.array_equal(offsets, eoff)
        assert np.array_equal(payload, epay)
        assert c.nbytes < 0.6 * ncb * sw * 4  # codewords as 16 bits, no fill
        d_back.fill_bytes(0x5A)
        bpc.compact_unpack(c, d_back, sw)
        assert bpc.status_take() == 0
        assert np.array_equal(d_back.download(), ostream)
        bpc.decode_device(d_back, d_rec, size, size, lut, p)
        assert bpc.status_take() == 0
        assert np.array_equal(d_rec.download().reshape(size, size), img)
        c.free()
    finally:
        for b in (d_img, d_cs, d_ncw, d_back, d_rec):
            b.free()
        lut.free()


def test_unpack_rejects_short_regions(bpc, O):
    """Codewords that do not fit the target regions: overflow status, MSB kept, -1 fill."""
    img, p, lut_arr, ostream, oncw = encoded(bpc, O, 512, 0, 15, 160)
    ncb, sw = p.n_codeblocks(512, 512), p.stride()
    d_cs = bpc.DeviceBuffer.from_array(ostream)
    c = bpc.compact_pack(d_cs, ncb, sw)
    assert bpc.status_take() == 0
    small = 64  # every coded codeblock here has more than 63 codewords
    assert (oncw[oncw > 0] > small - 1).all()
    d_small = bpc.DeviceBuffer(ncb * small)
    bpc.compact_unpack(c, d_small, small)
    assert bpc.status_take() == bpc.ERR_OVERFLOW
    got = d_small.download().reshape(ncb, small)
    assert np.array_equal(got[:, 0], ostream.reshape(ncb, sw)[:, 0])
    assert (got[:, 1:] == -1).all()
    c.free()
    d_cs.free()
    d_small.free()


def test_malformed_msb_flagged(bpc):
    """A region whose first word is not an MSB (0..27 or 32): byte 255 and malformed status."""
    sw = 4096
    regions = np.full((2, sw), -1, np.int32)
    regions[0, 0] = 5
    regions[0, 1:4] = [10, 20, 30]
    regions[1, 0] = 40
    d = bpc.DeviceBuffer.from_array(regions.ravel())
    c = bpc.compact_pack(d, 2, sw)
    assert bpc.status_take() == bpc.ERR_MALFORMED
    msb, offsets, payload = c.download()
    assert list(msb) == [5, 255] and list(offsets) == [0, 3, 3] and list(payload) == [10, 20, 30]
    c.free()
    d.free()


def test_bin_files_from_device(bpc, O, tmp_path):
    img, p, lut_arr, ostream, oncw = encoded(bpc, O, 2048, 1, 15, 170)
    ncb, sw = p.n_codeblocks(2048, 2048), p.stride()
    d_cs = bpc.DeviceBuffer.from_array(ostream)
    c = bpc.compact_pack(d_cs, ncb, sw)
    assert bpc.status_take() == 0
    msb, offsets, payload = c.download()
    bpc.bin_export(str(tmp_path), msb, offsets, payload, 2048, 2048, p)
    assert np.array_equal(bpc.bin_import(str(tmp_path), 2048, 2048, p), ostream)
    c.free()
    d_cs.free()


@pytest.mark.parametrize("size,cp3,nbp,rows,stride,layout", [
    (2048, 0, 15, 64, 0, 0), (1024, 1, 16, 32, 0, 1), (1024, 1, 15, 64, 6002, 0), (4096, 1, 16, 64, 0, 1),
    (256, 0, 15, 64, 0, 0), (1024, 1, 20, 64, 8192, 1), (512, 0, 24, 64, 12288, 0)])
def test_decode_compact_matches_oracle(bpc, O, size, cp3, nbp, rows, stride, layout):
    """Decoding straight from the compact form of the oracle's codestream gives the
    original coefficients (the per-image launch with and without register planes:
    256^2 and 1024^2 fit every wave at once, 4096^2 grid-strides over slots; 20 and
    24 LUT bitplanes take the 20- and 28-plane kernels, with regions wide enough)."""
    print("decode_compact", size, cp3, nbp, rows, stride, layout, flush=True)
    img, p, lut_arr, ostream, oncw = encoded(bpc, O, size, cp3, nbp, 170 + size // 256 + cp3, rows, stride,
                                             mag_bits=min(nbp, 22) if nbp >= 16 else 14,
                                             ll_bits=min(nbp, 22) - 1 if nbp > 16 else 13)
    p.layout = layout
    ncb, sw = p.n_codeblocks(size, size), p.stride()
    lut = bpc.Lut.from_array(lut_arr)
    d_cs = bpc.DeviceBuffer.from_array(ostream)
    d_rec = bpc.DeviceBuffer(size * size)
    try:
        c = bpc.compact_pack(d_cs, ncb, sw)
        assert bpc.status_take() == 0
        d_rec.fill_bytes(0x77)
        bpc.decode_compact_device(c, d_rec, size, size, lut, p)
        assert bpc.status_take() == 0
        got = d_rec.download()
        want = synth.to_tiled(img, rows).ravel() if layout else img.ravel()
        assert np.array_equal(got, want)
        c.free()
    finally:
        d_cs.free()
        d_rec.free()
        lut.free()


def test_decode_compact_equals_unpack_then_decode_on_damaged_input(bpc, O):
    """Damaged containers (a malformed MSB byte, counts past the region, a payload
    shorter than the offsets say, decreasing offsets) decode exactly as the
    regions that unpack makes of them: same coefficients, same status."""
    size, rows = 512, 64
    img, p, lut_arr, ostream, oncw = encoded(bpc, O, size, 1, 15, 180, rows)
    ncb, sw = p.n_codeblocks(size, size), p.stride()
    lut = bpc.Lut.from_array(lut_arr)
    d_cs = bpc.DeviceBuffer.from_array(ostream)
    c = bpc.compact_pack(d_cs, ncb, sw)
    assert bpc.status_take() == 0
    msb, offsets, payload = c.download()
    rng = np.random.default_rng(5)
    cases = []
    m = msb.copy(); m[3] = 255; cases.append(("malformed msb", m, offsets, payload, c.total))
    o = offsets.copy(); o[6:] += sw; cases.append(("count past region", msb, o, payload, c.total))
    cases.append(("short payload", msb, offsets, payload, c.total - 500))
    o = offsets.copy(); o[9] = o[11]; cases.append(("decreasing offsets", msb, o, payload, c.total))
    pl = payload.copy(); idx = rng.integers(0, c.total, 40); pl[idx] ^= 0x5A5A
    cases.append(("flipped codewords", msb, offsets, pl, c.total))
    d_regions = bpc.DeviceBuffer(ncb * sw)
    d_a = bpc.DeviceBuffer(size * size)
    d_b = bpc.DeviceBuffer(size * size)
    try:
        for name, m, o, pl, total in cases:
            print("damaged:", name, flush=True)
            dc = bpc.Compact(bpc.DeviceBuffer.from_array(m), bpc.DeviceBuffer.from_array(o),
                             bpc.DeviceBuffer.from_array(pl), ncb, int(total))
            try:
                bpc.compact_unpack(dc, d_regions, sw)
                st_unpack = bpc.status_take()
                bpc.decode_device(d_regions, d_a, size, size, lut, p)
                st_a = bpc.status_take()
                bpc.decode_compact_device(dc, d_b, size, size, lut, p)
                st_b = bpc.status_take()
                assert np.array_equal(d_a.download(), d_b.download()), name
                # one status word for both steps: the first of their codes in the
                # order bpc_status_take reports them
                order = [bpc.ERR_MALFORMED, bpc.ERR_RANGE, bpc.ERR_OVERFLOW, bpc.ERR_WIDE, 0]
                want = min(st_unpack, st_a, key=order.index)
                assert st_b == want, (name, st_unpack, st_a, st_b)
                assert name == "flipped codewords" or want != 0, name  # every other case is damaged
            finally:
                dc.free()
    finally:
        for b in (d_cs, d_regions, d_a, d_b):
            b.free()
        c.free()
        lut.free()


def test_decode_compact_all_zero_image_without_payload(bpc):
    """An all-zero image has MSB 32 everywhere and no codewords: its container has
    an empty payload, which may be passed as a null pointer; it decodes to zeros."""
    size = 512
    p = bpc.params(cpass3=1, lut_n_bitplanes=15)
    ncb, sw = p.n_codeblocks(size, size), p.stride()
    lut = bpc.Lut.from_array(np.full(bpc.lut_len(p), 64, np.int32))
    d_img = bpc.DeviceBuffer.from_array(np.zeros(size * size, np.int32))
    d_cs = bpc.DeviceBuffer(ncb * sw)
    d_rec = bpc.DeviceBuffer(size * size)
    try:
        bpc.encode_device(d_img, d_cs, size, size, lut, p)
        assert bpc.status_take() == 0
        c = bpc.compact_pack(d_cs, ncb, sw)
        assert bpc.status_take() == 0 and c.total == 0
        msb, offsets, _ = c.download()
        assert (msb == 32).all() and (offsets == 0).all()
        d_rec.fill_bytes(0x3C)
        empty = bpc.Compact(c.msb, c.offsets, None, ncb, 0)
        bpc.decode_compact_device(empty, d_rec, size, size, lut, p)
        assert bpc.status_take() == 0
        assert (d_rec.download() == 0).all()
        c.free()
    finally:
        for b in (d_img, d_cs, d_rec):
            b.free()
        lut.free()
