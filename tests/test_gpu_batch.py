"""The batched persistent path: every image of a mixed batch (uneven sizes, 2- and
3-pass, different LUTs and bitplane counts) must come out identical to coding it
alone with the per-image launch, and bit-exact against the oracle; graph replays
must give the same result every time."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def make_set(bpc, specs, seed, mag=None, stride=0, ll_bits=13, detail_exp=9):
    """specs: list of (size, cpass3, nbp, layout, rows); returns per-image dicts with device buffers.
    mag: magnitude bitplanes of the data (default min(nbp, 16)); stride: region words (0: reference)."""
    out = []
    for k, (size, cp3, nbp, layout, rows) in enumerate(specs):
        levels = 5 if size >= 1024 else 3
        mb = mag if mag is not None else min(nbp, 16)
        gen = dict(mag_bits=mb, ll_bits=min(ll_bits, mb), detail_exp=detail_exp)
        img = synth.synth_image(size, size, levels, seed + 2 * k, **gen)
        p = bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp, layout=layout,
                       cb_stride_words=stride)
        lut_arr = bpc.train_lut(synth.synth_image(size, size, levels, seed + 2 * k + 1, **gen),
                                bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp))
        src = img if layout == 0 else synth.to_tiled(img, rows)
        ncb = p.n_codeblocks(size, size)
        out.append(dict(img=img, p=p, lut_arr=lut_arr, lut=bpc.Lut.from_array(lut_arr),
                        d_in=bpc.DeviceBuffer.from_array(np.ascontiguousarray(src).ravel()),
                        d_cs=bpc.DeviceBuffer(bpc.stream_words(size, size, p)),
                        d_out=bpc.DeviceBuffer(size * size), d_ncw=bpc.DeviceBuffer(ncb), size=size))
    return out


def free_set(items):
    for it in items:
        for k in ("d_in", "d_cs", "d_out", "d_ncw"):
            it[k].free()
        it["lut"].free()


def plain_reference(bpc, it):
    """Codestream, ncw and reconstruction from the per-image launch."""
    s = it["size"]
    d_cs = bpc.DeviceBuffer(bpc.stream_words(s, s, it["p"]))
    d_ncw = bpc.DeviceBuffer(it["p"].n_codeblocks(s, s))
    d_out = bpc.DeviceBuffer(s * s)
    bpc.encode_device(it["d_in"], d_cs, s, s, it["lut"], it["p"], 1, d_ncw)
    bpc.decode_device(d_cs, d_out, s, s, it["lut"], it["p"])
    assert bpc.status_take() == 0
    r = (d_cs.download(), d_ncw.download(), d_out.download())
    for b in (d_cs, d_ncw, d_out):
        b.free()
    return r


def batch_encode_decode(bpc, batch, items, use_graph=False):
    enc = [bpc.make_image(it["d_in"], it["d_cs"], it["lut"], it["size"], it["size"], it["p"], it["d_ncw"])
           for it in items]
    batch.run(bpc.CODE, enc, use_graph=use_graph)
    dec = [bpc.make_image(it["d_cs"], it["d_out"], it["lut"], it["size"], it["size"], it["p"]) for it in items]
    batch.run(bpc.DECODE, dec, use_graph=use_graph)
    assert bpc.status_take() == 0, bpc.lib().bpc_last_error()


def check_oracle(O, items, k=None):
    """Codestream and codeword counts of every image against the oracle; the
    reconstruction against the original."""
    for j, it in enumerate(items):
        if k is not None and j != k:
            continue
        p, img = it["p"], it["img"]
        op = O.params(cpass3=p.cpass3, dwt_levels=p.dwt_levels, cblock_length=p.cblock_length,
                      lut_n_bitplanes=p.lut_n_bitplanes, cb_stride_words=p.cb_stride_words)
        st, ostream, oncw = O.encode(img, it["lut_arr"], op)
        assert st == 0, j
        assert np.array_equal(it["d_cs"].download(), ostream), j
        assert np.array_equal(it["d_ncw"].download(), oncw), j
        got = it["d_out"].download()
        got = got.reshape(img.shape) if p.layout == 0 else synth.from_tiled(got, it["size"], it["size"],
                                                                             p.cblock_length)
        assert np.array_equal(got, img), j


SPECS = [(512, 0, 15, 1, 64), (768, 1, 15, 0, 64), (1024, 1, 16, 1, 64), (512, 1, 15, 0, 64),
         (1024, 0, 14, 0, 64), (768, 0, 15, 1, 64), (192, 1, 15, 0, 64)]


@pytest.mark.parametrize("wpb,wps,cap", [(4, 0, 0), (1, 2, 0), (2, 3, 0), (4, 0, 5), (1, 0, 5), (2, 0, 6), (1, 0, 7), (4, 0, 8)])
def test_batch_matches_plain_and_oracle(bpc, O, wpb, wps, cap):
    items = make_set(bpc, SPECS, seed=1000 + wpb + cap)
    batch = bpc.Batch(16, 1 << 16)
    batch.set_tuning(wpb, wps)
    batch.set_vgpr_cap(cap)
    try:
        batch_encode_decode(bpc, batch, items)
        for k, it in enumerate(items):
            cs, ncw, rec = plain_reference(bpc, it)
            assert np.array_equal(it["d_cs"].download(), cs), k
            assert np.array_equal(it["d_ncw"].download(), ncw), k
            assert np.array_equal(it["d_out"].download(), rec), k
        check_oracle(O, items)
    finally:
        batch.destroy()
        free_set(items)


def test_batch_graph_replays(bpc):
    """The captured graph is replayed several times, with eager work and host reads
    in between; every replay must produce the eager result."""
    items = make_set(bpc, SPECS[:4], seed=2000)
    batch = bpc.Batch(8, 1 << 15)
    try:
        batch_encode_decode(bpc, batch, items, use_graph=False)
        want = [(it["d_cs"].download(), it["d_out"].download()) for it in items]
        for rep in range(4):
            for it in items:
                it["d_cs"].fill_bytes(0x33)
                it["d_out"].fill_bytes(0x44)
            bpc.device_sync()
            batch_encode_decode(bpc, batch, items, use_graph=True)
            for k, it in enumerate(items):
                assert np.array_equal(it["d_cs"].download(), want[k][0]), (rep, k)
                assert np.array_equal(it["d_out"].download(), want[k][1]), (rep, k)
            plain_reference(bpc, items[rep % len(items)])  # eager work between replays
    finally:
        batch.destroy()
        free_set(items)


def test_batch_alone_vs_in_mix(bpc):
    """A codeblock's output does not depend on what else is in the batch."""
    items = make_set(bpc, SPECS, seed=3000)
    batch = bpc.Batch(16, 1 << 16)
    try:
        alone = []
        for it in items:
            batch_encode_decode(bpc, batch, [it])
            alone.append((it["d_cs"].download(), it["d_out"].download()))
        batch_encode_decode(bpc, batch, items[::-1])
        for k, it in enumerate(items):
            assert np.array_equal(it["d_cs"].download(), alone[k][0]), k
            assert np.array_equal(it["d_out"].download(), alone[k][1]), k
    finally:
        batch.destroy()
        free_set(items)


def test_batch_rejects_mixed_codeblock_lengths(bpc):
    items = make_set(bpc, [(512, 0, 15, 0, 64), (512, 0, 15, 0, 32)], seed=4000)
    batch = bpc.Batch(4, 1 << 12)
    try:
        enc = [bpc.make_image(it["d_in"], it["d_cs"], it["lut"], it["size"], it["size"], it["p"]) for it in items]
        with pytest.raises(bpc.BPCError) as e:
            batch.run(bpc.CODE, enc)
        assert e.value.code == bpc.ERR_ARG
    finally:
        batch.destroy()
        free_set(items)


@pytest.mark.parametrize("rows,stride", [(32, 0), (16, 0), (2, 512)])
def test_batch_short_codeblocks_vs_oracle(bpc, O, rows, stride):
    """Codeblocks of 32, 16 and 2 rows (2 rows with regions large enough for 16-bit
    data), every image against the oracle, in both layouts."""
    specs = [(512, 0, 15, 1, rows), (256, 1, 16, 0, rows), (768, 1, 15, 1, rows), (320, 0, 15, 0, rows)]
    items = make_set(bpc, specs, seed=6000 + rows, stride=stride)
    batch = bpc.Batch(8, 1 << 16)
    try:
        batch_encode_decode(bpc, batch, items)
        check_oracle(O, items)
    finally:
        batch.destroy()
        free_set(items)


@pytest.mark.parametrize("nbp,mag,stride", [(20, 19, 8192), (28, 27, 16384)])
def test_batch_20_and_28_planes_vs_oracle(bpc, O, nbp, mag, stride):
    """The 20- and 28-plane kernels in the batch (wide magnitudes, regions larger
    than the reference's), every image against the oracle, at two register caps."""
    specs = [(256, 1, nbp, 1, 64), (512, 0, nbp, 0, 64), (320, 1, nbp, 0, 64)]
    items = make_set(bpc, specs, seed=7000 + nbp, mag=mag, stride=stride, ll_bits=mag, detail_exp=mag - 4)
    for cap in (0, 8):
        batch = bpc.Batch(4, 1 << 14)
        batch.set_vgpr_cap(cap)
        try:
            for it in items:
                it["d_cs"].fill_bytes(0x5A)
            batch_encode_decode(bpc, batch, items)
            check_oracle(O, items)
        finally:
            batch.destroy()
    free_set(items)


def test_batch_wide_codeblocks_vs_oracle(bpc, O):
    """Codeblocks whose MSB is beyond the planes the LUTs' bitplane count provides
    for (17-bit magnitudes, 15 LUT bitplanes; legal: the reference addresses such
    rows into the next LUT group) are coded by the batch's 28-plane pass, exactly as
    by the oracle, with no error status."""
    specs = [(512, 0, 15, 1, 64), (512, 1, 15, 0, 64), (768, 1, 15, 1, 64)]
    items = make_set(bpc, specs, seed=7500)
    for it in items:
        img = it["img"]
        s = it["size"]
        # finest-level HH band (LUT group 2: its rows at MSB 16 stay inside the table)
        img[s // 2 + 5, s // 2 + 9] = -(1 << 16) - 3
        img[s - 1, s - 64] = (1 << 16) + 77
        src = img if it["p"].layout == 0 else synth.to_tiled(img)
        it["d_in"].upload(np.ascontiguousarray(src).ravel())
    batch = bpc.Batch(4, 1 << 14)
    try:
        for graph in (False, True, True):
            batch_encode_decode(bpc, batch, items, use_graph=graph)
            check_oracle(O, items)
    finally:
        batch.destroy()
        free_set(items)


def test_batch_one_graph_serves_every_batch_shape(bpc, O):
    """The captured sequence reads the run's shape on the device: batches of
    different image counts, sizes, pass modes and codeblock lengths (64 and 32
    rows) replay one graph per direction, on a grid held at its cap (1 wave per
    workgroup, 1 wave per SIMD), and each comes out as the oracle's."""
    batch = bpc.Batch(4, 1 << 14)
    batch.set_tuning(1, 1)
    shapes = [
        [(4096, 1, 15, 1, 64)],
        [(1024, 0, 15, 1, 32), (512, 1, 15, 1, 32), (1536, 0, 15, 1, 32)],
        [(512, 0, 15, 1, 64), (2048, 1, 15, 1, 64)],
        [(4096, 0, 15, 1, 32)],
    ]
    try:
        for k, specs in enumerate(shapes):
            items = make_set(bpc, specs, seed=8000 + 10 * k, stride=4096)
            try:
                batch_encode_decode(bpc, batch, items, use_graph=True)
                check_oracle(O, items)
            finally:
                free_set(items)
            assert batch.graph_count() == 2  # encode, decode
    finally:
        batch.destroy()


def test_batch_graph_cache_is_bounded(bpc):
    """At most 32 captured graphs per batch object (one per direction, bitplane
    count and launch tuning); the least recently used go, and a batch replayed
    after its graph was dropped is captured again."""
    batch = bpc.Batch(2, 1 << 12)
    items = make_set(bpc, [(256, 0, 15, 1, 64)], seed=8500)
    it = items[0]
    tunings = [(wpb, wps) for wps in range(0, 9) for wpb in (1, 2, 4)]  # 27 x 2 directions
    try:
        for wpb, wps in tunings + tunings[:3]:
            batch.set_tuning(wpb, wps)
            it["d_out"].fill_bytes(0x55)
            batch_encode_decode(bpc, batch, [it], use_graph=True)
            assert batch.graph_count() <= 32
            got = synth.from_tiled(it["d_out"].download(), it["size"], it["size"])
            assert np.array_equal(got, it["img"]), (wpb, wps)
        assert batch.graph_count() == 32
    finally:
        batch.destroy()
        free_set(items)


def test_batch_graph_replay_folds_into_the_current_status_word(bpc):
    """A replayed graph reports into the status word of the current call: a
    malformed region decoded with a second status word marks that word only."""
    items = make_set(bpc, [(512, 1, 15, 1, 64)], seed=8700)
    it = items[0]
    st1 = bpc.DeviceBuffer(1, np.uint32)
    st2 = bpc.DeviceBuffer(1, np.uint32)
    st1.fill_bytes(0)
    st2.fill_bytes(0)
    batch = bpc.Batch(2, 1 << 10)
    try:
        enc = [bpc.make_image(it["d_in"], it["d_cs"], it["lut"], 512, 512, it["p"], it["d_ncw"])]
        dec = [bpc.make_image(it["d_cs"], it["d_out"], it["lut"], 512, 512, it["p"])]
        batch.run(bpc.CODE, enc, use_graph=True, d_status=st1)
        batch.run(bpc.DECODE, dec, use_graph=True, d_status=st1)
        assert bpc.status_take(st1) == 0
        cs = it["d_cs"].download()
        cs[0] = 40  # MSB word outside [0, 27]
        it["d_cs"].upload(cs)
        batch.run(bpc.DECODE, dec, use_graph=True, d_status=st2)
        assert bpc.status_take(st1) == 0
        assert bpc.status_take(st2) == bpc.ERR_MALFORMED
        assert batch.graph_count() == 2
    finally:
        batch.destroy()
        free_set(items)
        st1.free()
        st2.free()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_int16_and_int32_images_vs_oracle(bpc, O, graph):
    """A batch mixing int16 and int32 coefficient images (raster and tiled, 2- and
    3-pass): every codestream and count as the oracle's, every int16 reconstruction
    equal to the original."""
    specs = [(512, 0, 15, 1, 64), (768, 1, 15, 0, 64), (1024, 1, 16, 1, 64), (512, 1, 15, 0, 64), (256, 0, 14, 1, 64)]
    items = make_set(bpc, specs, seed=9400, mag=14)
    wide16 = {0, 2, 4}
    buf16 = {}
    try:
        for k in wide16:
            it = items[k]
            src = it["img"] if it["p"].layout == 0 else synth.to_tiled(it["img"], it["p"].cblock_length)
            buf16[k] = (bpc.DeviceBuffer.from_array(np.ascontiguousarray(src).astype(np.int16).ravel()),
                        bpc.DeviceBuffer(it["size"] ** 2, np.int16))
        batch = bpc.Batch(8, 1 << 14)
        try:
            enc = [bpc.make_image(buf16[k][0] if k in buf16 else it["d_in"], it["d_cs"], it["lut"], it["size"],
                                  it["size"], it["p"], it["d_ncw"], int16=k in buf16) for k, it in enumerate(items)]
            dec = [bpc.make_image(it["d_cs"], buf16[k][1] if k in buf16 else it["d_out"], it["lut"], it["size"],
                                  it["size"], it["p"], int16=k in buf16) for k, it in enumerate(items)]
            for it in items:
                it["d_cs"].fill_bytes(0x21)
            batch.run(bpc.CODE, enc, use_graph=graph)
            batch.run(bpc.DECODE, dec, use_graph=graph)
            assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
        finally:
            batch.destroy()
        for k, it in enumerate(items):
            p, img = it["p"], it["img"]
            op = O.params(cpass3=p.cpass3, dwt_levels=p.dwt_levels, cblock_length=p.cblock_length,
                          lut_n_bitplanes=p.lut_n_bitplanes)
            st, ostream, oncw = O.encode(img, it["lut_arr"], op)
            assert st == 0
            assert np.array_equal(it["d_cs"].download(), ostream), k
            assert np.array_equal(it["d_ncw"].download(), oncw), k
            got = (buf16[k][1] if k in buf16 else it["d_out"]).download().astype(np.int32)
            got = got.reshape(img.shape) if p.layout == 0 else synth.from_tiled(got, it["size"], it["size"])
            assert np.array_equal(got, img), k
    finally:
        for a, b in buf16.values():
            a.free()
            b.free()
        free_set(items)


def test_batch_int16_decode_saturates(bpc, O):
    """In a batch, an int16 decode of magnitudes past the int16 range saturates and
    reports BPC_ERR_RANGE, while an int32 image of the same batch decodes exactly."""
    size = 256
    img = synth.synth_image(size, size, 3, 211, ll_bits=17, mag_bits=17)
    assert (np.abs(img.astype(np.int64)) > 32767).any()
    p = bpc.params(cpass3=1, dwt_levels=3, lut_n_bitplanes=18)
    lut_arr = bpc.train_lut(img, p)
    st, ostream, _ = O.encode(img, lut_arr, O.params(cpass3=1, dwt_levels=3, lut_n_bitplanes=18))
    assert st == 0
    lut = bpc.Lut.from_array(lut_arr)
    cs = bpc.DeviceBuffer.from_array(ostream)
    out16 = bpc.DeviceBuffer(size * size, np.int16)
    out32 = bpc.DeviceBuffer(size * size)
    batch = bpc.Batch(2, 1 << 10)
    try:
        dec = [bpc.make_image(cs, out16, lut, size, size, p, int16=True), bpc.make_image(cs, out32, lut, size, size, p)]
        batch.run(bpc.DECODE, dec)
        assert bpc.status_take() == bpc.ERR_RANGE
        assert np.array_equal(out16.download().reshape(size, size), np.clip(img, -32768, 32767).astype(np.int16))
        assert np.array_equal(out32.download().reshape(size, size), img)
    finally:
        batch.destroy()
        for b in (cs, out16, out32):
            b.free()
        lut.free()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_at_image_capacity_vs_oracle(bpc, O, graph):
    """4096 small images in one batch (the capacity): the staging area read by the
    run's first kernel at its largest, one- and two-codeblock images (odd counts
    leave padding entries), 2- and 3-pass alternating; every codestream and
    reconstruction against the oracle."""
    n = bpc.MAX_BATCH_IMAGES if hasattr(bpc, "MAX_BATCH_IMAGES") else 4096
    rng = np.random.default_rng(77)
    shapes = [(64, 64), (128, 64)]
    luts = {}
    for cp3 in (0, 1):
        train = synth.synth_image(128, 64, 1, 7700 + cp3, mag_bits=12)
        luts[cp3] = bpc.train_lut(train, bpc.params(cpass3=cp3, dwt_levels=1, lut_n_bitplanes=13))
    lut_dev = {cp3: bpc.Lut.from_array(a) for cp3, a in luts.items()}
    items = []
    try:
        for k in range(n):
            sx, sy = shapes[k % 2]
            cp3 = (k // 2) % 2
            img = synth.synth_image(sx, sy, 1, 8000 + k, mag_bits=12) if k % 7 else np.zeros((sy, sx), np.int32)
            p = bpc.params(cpass3=cp3, dwt_levels=1, lut_n_bitplanes=13)
            ncb = p.n_codeblocks(sx, sy)
            items.append(dict(img=img, sx=sx, sy=sy, cp3=cp3, p=p, d_in=bpc.DeviceBuffer.from_array(img.ravel()),
                              d_cs=bpc.DeviceBuffer(bpc.stream_words(sx, sy, p)), d_out=bpc.DeviceBuffer(sx * sy),
                              d_ncw=bpc.DeviceBuffer(ncb)))
        total = sum(it["p"].n_codeblocks(it["sx"], it["sy"]) for it in items)
        batch = bpc.Batch(n, total)
        try:
            enc = [bpc.make_image(it["d_in"], it["d_cs"], lut_dev[it["cp3"]], it["sx"], it["sy"], it["p"], it["d_ncw"])
                   for it in items]
            dec = [bpc.make_image(it["d_cs"], it["d_out"], lut_dev[it["cp3"]], it["sx"], it["sy"], it["p"])
                   for it in items]
            batch.run(bpc.CODE, enc, use_graph=graph)
            batch.run(bpc.DECODE, dec, use_graph=graph)
            assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
        finally:
            batch.destroy()
        for k, it in enumerate(items):
            st, ostream, oncw = O.encode(it["img"], luts[it["cp3"]],
                                         O.params(cpass3=it["cp3"], dwt_levels=1, lut_n_bitplanes=13))
            assert st == 0
            assert np.array_equal(it["d_cs"].download(), ostream), k
            assert np.array_equal(it["d_ncw"].download(), oncw), k
            assert np.array_equal(it["d_out"].download().reshape(it["sy"], it["sx"]), it["img"]), k
    finally:
        for it in items:
            for key in ("d_in", "d_cs", "d_out", "d_ncw"):
                it[key].free()
        for l in lut_dev.values():
            l.free()
