"""The batched persistent path: every image of a mixed batch (uneven sizes, 2- and
3-pass, different LUTs and bitplane counts) must come out identical to coding it
alone with the per-image launch, and bit-exact against the oracle; graph replays
must give the same result every time."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def make_set(bpc, specs, seed):
    """specs: list of (size, cpass3, nbp, layout, rows); returns per-image dicts with device buffers."""
    out = []
    for k, (size, cp3, nbp, layout, rows) in enumerate(specs):
        levels = 5 if size >= 1024 else 3
        img = synth.synth_image(size, size, levels, seed + 2 * k, mag_bits=min(nbp, 16))
        p = bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp, layout=layout)
        lut_arr = bpc.train_lut(synth.synth_image(size, size, levels, seed + 2 * k + 1, mag_bits=min(nbp, 16)),
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
            img = it["img"]
            got = it["d_out"].download()
            got = got.reshape(img.shape) if it["p"].layout == 0 else synth.from_tiled(got, it["size"], it["size"])
            assert np.array_equal(got, img), k
            if k < 3:
                p = it["p"]
                st, ostream, _ = O.encode(img, it["lut_arr"], O.params(cpass3=p.cpass3, dwt_levels=p.dwt_levels,
                                                                       lut_n_bitplanes=p.lut_n_bitplanes))
                assert st == 0 and np.array_equal(it["d_cs"].download(), ostream), k
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
