"""bpc_codec_device_noalloc: the per-image codec with its bitplane scratch in a fixed
per-device array, for callers that cannot allocate.  Its results must be those of
bpc_encode_device / bpc_decode_device (bit-exact), on two calls in a row with the
second on a smaller image (the array keeps the first call's contents), for the 16-
and 28-plane kernels, and the calls must not change the device's free memory."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def make(bpc, size, cp3, nbp, layout, seed, mag=None, stride=0):
    levels = 5 if size >= 1024 else 3
    mb = mag if mag is not None else min(nbp, 16)
    gen = dict(mag_bits=mb, ll_bits=min(13, mb) if mag is None else mb, detail_exp=9 if mag is None else mb - 4)
    img = synth.synth_image(size, size, levels, seed, **gen)
    lut = bpc.train_lut(synth.synth_image(size, size, levels, seed + 1, **gen),
                        bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp))
    p = bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp, layout=layout, cb_stride_words=stride)
    src = img if layout == bpc.LAYOUT_RASTER else synth.to_tiled(img)
    return dict(size=size, p=p, lut=bpc.Lut.from_array(lut), d_img=bpc.DeviceBuffer.from_array(
        np.ascontiguousarray(src).ravel()), src=np.ascontiguousarray(src).ravel())


def reference(bpc, it):
    s, p = it["size"], it["p"]
    d_cs = bpc.DeviceBuffer(bpc.stream_words(s, s, p))
    d_ncw = bpc.DeviceBuffer(p.n_codeblocks(s, s))
    d_out = bpc.DeviceBuffer(s * s)
    bpc.encode_device(it["d_img"], d_cs, s, s, it["lut"], p, 1, d_ncw)
    bpc.decode_device(d_cs, d_out, s, s, it["lut"], p)
    assert bpc.status_take() == 0
    out = (d_cs.download(), d_ncw.download(), d_out.download())
    for b in (d_cs, d_ncw, d_out):
        b.free()
    return out


def noalloc(bpc, it, d_cs, d_ncw, d_out, d_status, stream):
    s, p = it["size"], it["p"]
    bpc.codec_device_noalloc(bpc.CODE, it["d_img"], d_cs, s, s, it["lut"], p, d_status, 1, d_ncw, stream)
    bpc.codec_device_noalloc(bpc.DECODE, d_cs, d_out, s, s, it["lut"], p, d_status, stream=stream)
    assert bpc.status_take(d_status, stream) == 0, bpc.lib().bpc_last_error()
    n_cs = bpc.stream_words(s, s, p)
    return d_cs.download(n_cs, stream), d_ncw.download(p.n_codeblocks(s, s), stream), d_out.download(s * s, stream)


@pytest.mark.parametrize("cases", [
    # (size, cpass3, LUT bitplanes, layout, magnitude bits, region words): a large
    # image, then a smaller one
    [(4096, 1, 16, 1, None, 0), (1024, 0, 15, 0, None, 0)],
    [(1024, 1, 28, 1, 27, 16384), (512, 0, 20, 0, 19, 8192)],
])
def test_noalloc_matches_device_entry_points(bpc, cases):
    items = [make(bpc, size, cp3, nbp, layout, 300 + 7 * k, mag, stride)
             for k, (size, cp3, nbp, layout, mag, stride) in enumerate(cases)]
    refs = [reference(bpc, it) for it in items]
    # buffers sized for the largest call, shared by both calls
    n_cs = max(bpc.stream_words(it["size"], it["size"], it["p"]) for it in items)
    n_img = max(it["size"] ** 2 for it in items)
    d_cs, d_out = bpc.DeviceBuffer(n_cs), bpc.DeviceBuffer(n_img)
    d_ncw = bpc.DeviceBuffer(max(it["p"].n_codeblocks(it["size"], it["size"]) for it in items))
    d_status = bpc.DeviceBuffer(1, np.uint32)
    d_status.fill_bytes(0)
    bpc.device_sync()  # the fill ran on the null stream; the calls below use their own
    stream = bpc.Stream()
    try:
        noalloc(bpc, items[0], d_cs, d_ncw, d_out, d_status, stream)  # first use on this device
        free_before, _ = bpc.mem_info()
        for _ in range(2):
            for it, (cs, ncw, rec) in zip(items, refs):
                got_cs, got_ncw, got_rec = noalloc(bpc, it, d_cs, d_ncw, d_out, d_status, stream)
                assert np.array_equal(got_cs, cs)
                assert np.array_equal(got_ncw, ncw)
                assert np.array_equal(got_rec, rec)
                assert np.array_equal(got_rec, it["src"])
        free_after, _ = bpc.mem_info()
        assert free_after == free_before
    finally:
        for b in (d_cs, d_out, d_ncw, d_status):
            b.free()
        stream.destroy()
        for it in items:
            it["d_img"].free()
            it["lut"].free()


def test_noalloc_requires_status_word(bpc):
    it = make(bpc, 256, 0, 15, 0, 5)
    d_cs = bpc.DeviceBuffer(bpc.stream_words(256, 256, it["p"]))
    try:
        with pytest.raises(bpc.BPCError) as e:
            bpc.codec_device_noalloc(bpc.CODE, it["d_img"], d_cs, 256, 256, it["lut"], it["p"], None)
        assert e.value.code == bpc.ERR_ARG
        with pytest.raises(bpc.BPCError):
            bpc.codec_device_noalloc(7, it["d_img"], d_cs, 256, 256, it["lut"], it["p"], d_cs)
    finally:
        d_cs.free()
        it["d_img"].free()
        it["lut"].free()
