"""The continuous-batching server: every request's output (encode: codestream
regions and codeword counts; decode: coefficients) must be identical whether it
ran alone (per-image launch) or in the mix, with requests arriving while others
are on the GPU, and with admission throttled by a small HBM budget."""
import random
import threading
import time

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def make_pool(bpc, specs, seed):
    pool = []
    for k, (size, cp3, nbp) in enumerate(specs):
        img = synth.synth_image(size, size, 5, seed + 2 * k, mag_bits=nbp - 1)
        p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
        lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, seed + 2 * k + 1, mag_bits=nbp - 1),
                                bpc.params(cpass3=cp3, lut_n_bitplanes=nbp))
        lut = bpc.Lut.from_array(lut_arr)
        d_img = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
        d_cs = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
        d_ncw = bpc.DeviceBuffer(p.n_codeblocks(size, size))
        d_out = bpc.DeviceBuffer(size * size)
        bpc.encode_device(d_img, d_cs, size, size, lut, p, 1, d_ncw)
        bpc.decode_device(d_cs, d_out, size, size, lut, p)
        assert bpc.status_take() == 0
        pool.append(dict(size=size, p=p, lut=lut, d_img=d_img, d_cs=d_cs, img=img,
                         cs=d_cs.download(), ncw=d_ncw.download(), rec=d_out.download()))
        d_ncw.free()
        d_out.free()
    return pool


def free_pool(pool):
    for it in pool:
        it["d_img"].free()
        it["d_cs"].free()
        it["lut"].free()


SPECS = [(512, 0, 15), (1024, 1, 16), (768, 1, 14), (1536, 0, 15), (512, 1, 16)]


@pytest.mark.parametrize("inflight,budget_mib", [(2, 4096), (3, 64), (1, 4096)])
def test_server_results_match_alone(bpc, inflight, budget_mib):
    pool = make_pool(bpc, SPECS, seed=5000 + inflight)
    srv = bpc.Server(max_inflight=inflight, max_batch_images=8, hbm_budget_bytes=budget_mib << 20)
    rng = random.Random(7)
    try:
        # open-loop arrivals from a second thread while the scheduler runs batches
        plan = [(rng.choice(range(len(pool))), rng.random() < 0.5) for _ in range(40)]
        ids = [None] * len(plan)

        def producer():
            for k, (j, enc) in enumerate(plan):
                it = pool[j]
                src = it["d_img"] if enc else it["d_cs"]
                ids[k] = srv.submit(bpc.CODE if enc else bpc.DECODE, src, it["lut"], it["size"], it["size"], it["p"])
                time.sleep(rng.random() * 0.002)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        lat = []
        for k, (j, enc) in enumerate(plan):
            res = srv.wait(ids[k], timeout_ms=60000)
            assert res.status == 0, k
            it = pool[j]
            if enc:
                got = bpc.download_ptr(res.output, it["cs"].size)
                assert np.array_equal(got, it["cs"]), k
                ncb = it["p"].n_codeblocks(it["size"], it["size"])
                assert np.array_equal(bpc.download_ptr(res.ncw, ncb), it["ncw"]), k
            else:
                got = bpc.download_ptr(res.output, it["size"] ** 2)
                assert np.array_equal(got, it["rec"]), k
            lat.append(res.latency_us)
            srv.release(ids[k])
        st = srv.stats()
        assert st.completed == len(plan) and st.pending == 0 and st.committed_bytes == 0
        assert min(lat) > 0
    finally:
        srv.destroy()
        free_pool(pool)


def test_server_rejects_oversized_request(bpc):
    srv = bpc.Server(hbm_budget_bytes=1 << 20)
    try:
        p = bpc.params()
        d = bpc.DeviceBuffer(1024 * 1024)
        lut = bpc.Lut.from_array(np.full(bpc.lut_len(p), 64, np.int32))
        with pytest.raises(bpc.BPCError):
            srv.submit(bpc.CODE, d, lut, 1024, 1024, p)  # 4 MiB output > 1 MiB budget
        d.free()
        lut.free()
    finally:
        srv.destroy()
