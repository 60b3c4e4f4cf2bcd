"""The continuous-batching server: every request's output (encode: codestream
regions and codeword counts; decode: coefficients) must equal the oracle's
whether it ran alone or in the mix, with requests arriving while others are on the
GPU, and with admission throttled by a small output budget or by the device's
free memory (an injected figure); malformed requests fail alone at submit."""
import random
import threading
import time

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def make_pool(bpc, O, specs, seed, wide=False):
    """Request images with their LUTs and the oracle's codestream (which the decode
    requests replay); wide: a few 17-bit coefficients beyond 15 LUT bitplanes."""
    pool = []
    for k, (size, cp3, nbp) in enumerate(specs):
        img = synth.synth_image(size, size, 5, seed + 2 * k, mag_bits=nbp - 1)
        if wide:  # finest-level HH codeblocks (LUT group 2), MSB 16
            img[size // 2 + 3, size // 2 + 70] = (1 << 16) + 9
            img[size - 2, size - 1] = -(1 << 16) - 1
        p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
        lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, seed + 2 * k + 1, mag_bits=nbp - 1),
                                bpc.params(cpass3=cp3, lut_n_bitplanes=nbp))
        st, cs, ncw = O.encode(img, lut_arr, O.params(cpass3=cp3, lut_n_bitplanes=nbp))
        assert st == 0
        lut = bpc.Lut.from_array(lut_arr)
        d_img = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
        d_cs = bpc.DeviceBuffer.from_array(cs)
        pool.append(dict(size=size, p=p, lut=lut, d_img=d_img, d_cs=d_cs, img=img, cs=cs, ncw=ncw,
                         rec=synth.to_tiled(img)))
    return pool


def check_result(bpc, it, enc, res):
    assert res.status == 0
    if enc:
        assert np.array_equal(bpc.download_ptr(res.output, it["cs"].size), it["cs"])
        ncb = it["p"].n_codeblocks(it["size"], it["size"])
        assert np.array_equal(bpc.download_ptr(res.ncw, ncb), it["ncw"])
    else:
        assert np.array_equal(bpc.download_ptr(res.output, it["size"] ** 2), it["rec"])


def free_pool(pool):
    for it in pool:
        it["d_img"].free()
        it["d_cs"].free()
        it["lut"].free()


SPECS = [(512, 0, 15), (1024, 1, 16), (768, 1, 14), (1536, 0, 15), (512, 1, 16)]


@pytest.mark.parametrize("inflight,budget_mib", [(2, 4096), (3, 64), (1, 4096)])
def test_server_results_match_oracle(bpc, O, inflight, budget_mib):
    pool = make_pool(bpc, O, SPECS, seed=5000 + inflight)
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
            check_result(bpc, pool[j], enc, res)
            lat.append(res.latency_us)
            srv.release(ids[k])
        st = srv.stats()
        assert st.completed == len(plan) and st.pending == 0 and st.committed_bytes == 0
        assert min(lat) > 0
    finally:
        srv.destroy()
        free_pool(pool)


def test_server_rejects_malformed_requests_alone(bpc, O):
    """Requests that could never run are refused at submit (they neither fail the
    batch they would join nor block the queue); a valid request still goes through."""
    pool = make_pool(bpc, O, [(512, 1, 15)], seed=5900)
    it = pool[0]
    srv = bpc.Server(max_inflight=1, max_batch_codeblocks=32)  # 512^2 = 64 codeblocks
    srv2 = bpc.Server(max_inflight=1)
    try:
        with pytest.raises(bpc.BPCError) as e:
            srv.submit(bpc.CODE, it["d_img"], it["lut"], 512, 512, it["p"])
        assert e.value.code == bpc.ERR_ARG and "codeblocks" in str(e.value)
        bad = bpc.params(cpass3=2)
        with pytest.raises(bpc.BPCError):
            srv2.submit(bpc.CODE, it["d_img"], it["lut"], 512, 512, bad)
        with pytest.raises(bpc.BPCError):
            srv2.submit(bpc.CODE, it["d_img"], it["lut"], 512 + 32, 512, it["p"])
        short = bpc.Lut(it["lut"].ptr, it["lut"].len - 1)
        with pytest.raises(bpc.BPCError):
            srv2.submit(bpc.CODE, it["d_img"], short, 512, 512, it["p"])
        with pytest.raises(bpc.BPCError):
            srv2.submit(bpc.CODE, None, it["lut"], 512, 512, it["p"])
        rid = srv2.submit(bpc.CODE, it["d_img"], it["lut"], 512, 512, it["p"])
        check_result(bpc, it, True, srv2.wait(rid, timeout_ms=60000))
        srv2.release(rid)
    finally:
        srv.destroy()
        srv2.destroy()
        free_pool(pool)


def test_server_admission_by_free_memory(bpc, O):
    """Admission follows the device's free memory, here an injected figure: with
    room for one output a second request waits (its wait times out), stays waiting
    after the first is released while the figure says memory is short, and runs as
    soon as the figure has room again.  Released outputs are not pooled here, so
    every admission needs free memory."""
    pool = make_pool(bpc, O, [(1024, 0, 15)], seed=5950)
    it = pool[0]
    out_bytes = 1024 * 1024 * 4  # decode output
    reserve = 64 << 20
    fake = {"free": reserve + out_bytes}
    srv = bpc.Server(max_inflight=1, hbm_reserve_bytes=reserve, pool_limit_bytes=0,
                     mem_info=lambda: (fake["free"], 288 << 30))
    try:
        a = srv.submit(bpc.DECODE, it["d_cs"], it["lut"], 1024, 1024, it["p"])
        check_result(bpc, it, False, srv.wait(a, timeout_ms=60000))
        fake["free"] = reserve + out_bytes - 1  # the first output holds the memory
        b = srv.submit(bpc.DECODE, it["d_cs"], it["lut"], 1024, 1024, it["p"])
        with pytest.raises(bpc.BPCError) as e:
            srv.wait(b, timeout_ms=300)
        assert e.value.code == bpc.ERR_TIMEOUT
        srv.release(a)
        with pytest.raises(bpc.BPCError):
            srv.wait(b, timeout_ms=300)
        assert srv.stats().pending == 1
        fake["free"] = reserve + out_bytes
        check_result(bpc, it, False, srv.wait(b, timeout_ms=60000))
        srv.release(b)
        st = srv.stats()
        assert st.completed == 2 and st.pending == 0 and st.committed_bytes == 0
    finally:
        srv.destroy()
        free_pool(pool)


def test_server_wide_codeblocks(bpc, O):
    """17-bit coefficients with 15 LUT bitplanes: the batch's 28-plane pass codes
    them, so the server returns the oracle's results with no error."""
    pool = make_pool(bpc, O, [(1024, 0, 15), (1024, 1, 15)], seed=5970, wide=True)
    srv = bpc.Server(max_inflight=2)
    try:
        plan = [(j, enc) for j in range(2) for enc in (True, False)] * 2
        ids = [srv.submit(bpc.CODE if enc else bpc.DECODE, pool[j]["d_img"] if enc else pool[j]["d_cs"],
                          pool[j]["lut"], 1024, 1024, pool[j]["p"]) for j, enc in plan]
        for rid, (j, enc) in zip(ids, plan):
            check_result(bpc, pool[j], enc, srv.wait(rid, timeout_ms=60000))
            srv.release(rid)
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


def test_server_int16_requests_vs_oracle(bpc, O):
    """int16 encode and decode requests served in the same batches as int32 ones:
    codestreams and counts as the oracle's, int16 outputs equal to the originals."""
    pool = make_pool(bpc, O, [(512, 0, 15), (1024, 1, 15), (768, 1, 14)], seed=5600)
    for it in pool:
        it["d_img16"] = bpc.DeviceBuffer.from_array(it["rec"].astype(np.int16).ravel())
    srv = bpc.Server(max_inflight=2, max_batch_images=8)
    rng = random.Random(3)
    plan = [(rng.randrange(len(pool)), rng.random() < 0.5, rng.random() < 0.6) for _ in range(30)]
    try:
        ids = []
        for j, enc, i16 in plan:
            it = pool[j]
            src = (it["d_img16"] if i16 else it["d_img"]) if enc else it["d_cs"]
            ids.append(srv.submit(bpc.CODE if enc else bpc.DECODE, src, it["lut"], it["size"], it["size"], it["p"],
                                  int16=i16))
        for rid, (j, enc, i16) in zip(ids, plan):
            it = pool[j]
            res = srv.wait(rid, timeout_ms=60000)
            assert res.status == 0
            if enc:
                assert np.array_equal(bpc.download_ptr(res.output, it["cs"].size), it["cs"])
                ncb = it["p"].n_codeblocks(it["size"], it["size"])
                assert np.array_equal(bpc.download_ptr(res.ncw, ncb), it["ncw"])
            else:
                got = bpc.download_ptr(res.output, it["size"] ** 2, np.int16 if i16 else np.int32)
                assert np.array_equal(got.astype(np.int32), it["rec"].ravel())
            srv.release(rid)
    finally:
        srv.destroy()
        for it in pool:
            it["d_img16"].free()
        free_pool(pool)
