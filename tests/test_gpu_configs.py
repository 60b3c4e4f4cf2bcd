"""The batched and served workloads at their own sizes, against the oracle.

Batched: 96 images in one persistent launch per direction, 32 each of 2048^2,
3072^2 and 4096^2, 5-level transform data of 14 magnitude bitplanes, 2- and 3-pass
alternating, a LUT per image (BPC_PaCo.cu:1097-1224 per image).  The images are
drawn from six distinct (size, pass mode) images, each with its own trained LUT,
so that the oracle codes each distinct image once; every one of the 96 outputs is
compared with it.

Served: a mix of encode and decode requests over {2048^2, 3072^2, 4096^2, 6144^2},
14-16 bitplanes, 2 or 3 passes, submitted from two threads while batches run;
every encode result against the oracle's codestream and counts, every decode
(of the oracle's codestream) against the original image."""
import random
import sys
import threading
import time

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def distinct_images(bpc, O, specs, seed):
    """specs: (size, cpass3, nbp, mag_bits); returns dicts with the image, its LUT
    and the oracle's codestream and counts (reference layout)."""
    out = []
    for k, (size, cp3, nbp, mag) in enumerate(specs):
        img = synth.synth_image(size, size, 5, seed + 2 * k, mag_bits=mag)
        lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, seed + 2 * k + 1, mag_bits=mag),
                                bpc.params(cpass3=cp3, lut_n_bitplanes=nbp))
        st, cs, ncw = O.encode(img, lut_arr, O.params(cpass3=cp3, lut_n_bitplanes=nbp))
        assert st == 0
        p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
        out.append(dict(size=size, p=p, img=img, cs=cs, ncw=ncw, lut=bpc.Lut.from_array(lut_arr),
                        d_img=bpc.DeviceBuffer.from_array(synth.to_tiled(img))))
        say("  image %d: %dx%d, %d passes, oracle done" % (k, size, size, 3 if cp3 else 2))
    return out


def test_batched_96_images_vs_oracle(bpc, O):
    sizes = (2048, 3072, 4096)
    say("batched_96: six distinct images, LUTs and oracle codestreams")
    base = distinct_images(bpc, O, [(s, cp3, 15, 14) for s in sizes for cp3 in (0, 1)], seed=9100)
    entries = []
    for k in range(96):
        b = base[2 * (k // 32) + (k % 2)]  # 32 per size, pass mode alternating
        s = b["size"]
        entries.append(dict(base=b, d_cs=bpc.DeviceBuffer(bpc.stream_words(s, s, b["p"])),
                            d_out=bpc.DeviceBuffer(s * s), d_ncw=bpc.DeviceBuffer(b["p"].n_codeblocks(s, s))))
    total_cbs = sum(e["base"]["p"].n_codeblocks(e["base"]["size"], e["base"]["size"]) for e in entries)
    assert total_cbs == 32 * (1024 + 2304 + 4096)
    batch = bpc.Batch(96, total_cbs)
    batch.set_vgpr_cap(8)
    try:
        enc = [bpc.make_image(e["base"]["d_img"], e["d_cs"], e["base"]["lut"], e["base"]["size"], e["base"]["size"],
                              e["base"]["p"], e["d_ncw"]) for e in entries]
        dec = [bpc.make_image(e["d_cs"], e["d_out"], e["base"]["lut"], e["base"]["size"], e["base"]["size"],
                              e["base"]["p"]) for e in entries]
        for graph in (False, True):
            say("batched_96: encode + decode, graph=%s" % graph)
            for e in entries:
                e["d_out"].fill_bytes(0x11)
            batch.run(bpc.CODE, enc, use_graph=graph)
            batch.run(bpc.DECODE, dec, use_graph=graph)
            assert bpc.status_take() == 0, bpc.lib().bpc_last_error()
            for k, e in enumerate(entries):
                b = e["base"]
                assert np.array_equal(e["d_cs"].download(), b["cs"]), k
                assert np.array_equal(e["d_ncw"].download(), b["ncw"]), k
                assert np.array_equal(synth.from_tiled(e["d_out"].download(), b["size"], b["size"]), b["img"]), k
    finally:
        batch.destroy()
        for e in entries:
            for key in ("d_cs", "d_out", "d_ncw"):
                e[key].free()
        for b in base:
            b["d_img"].free()
            b["lut"].free()


def test_served_mix_with_6144_vs_oracle(bpc, O):
    variants = [(0, 15, 14), (1, 16, 16)]  # (cpass3, LUT bitplanes, magnitude bitplanes)
    specs = [(s, cp3, nbp, mag) for s in (2048, 3072, 4096, 6144) for cp3, nbp, mag in variants]
    say("served: eight request images, LUTs and oracle codestreams")
    pool = distinct_images(bpc, O, specs, seed=9300)
    for it in pool:
        it["d_cs"] = bpc.DeviceBuffer.from_array(it["cs"])
    rng = random.Random(11)
    plan = [(rng.randrange(len(pool)), rng.random() < 0.5) for _ in range(48)]
    for j in range(len(pool)):  # every image, both directions, at least once
        plan += [(j, True), (j, False)]
    rng.shuffle(plan)
    srv = bpc.Server(max_inflight=3, max_batch_images=16)
    ids = [None] * len(plan)
    try:
        def producer(part):
            for k in range(part, len(plan), 2):
                j, enc = plan[k]
                it = pool[j]
                ids[k] = srv.submit(bpc.CODE if enc else bpc.DECODE, it["d_img"] if enc else it["d_cs"], it["lut"],
                                    it["size"], it["size"], it["p"])
                time.sleep(rng.random() * 0.003)

        say("served: %d requests from two threads" % len(plan))
        threads = [threading.Thread(target=producer, args=(t,)) for t in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for k, (j, enc) in enumerate(plan):
            res = srv.wait(ids[k], timeout_ms=120000)
            it = pool[j]
            assert res.status == 0, k
            if enc:
                assert np.array_equal(bpc.download_ptr(res.output, it["cs"].size), it["cs"]), k
                ncb = it["p"].n_codeblocks(it["size"], it["size"])
                assert np.array_equal(bpc.download_ptr(res.ncw, ncb), it["ncw"]), k
            else:
                got = bpc.download_ptr(res.output, it["size"] ** 2)
                assert np.array_equal(synth.from_tiled(got, it["size"], it["size"]), it["img"]), k
            srv.release(ids[k])
        st = srv.stats()
        assert st.completed == len(plan) and st.pending == 0 and st.committed_bytes == 0
    finally:
        srv.destroy()
        for it in pool:
            it["d_img"].free()
            it["d_cs"].free()
            it["lut"].free()
