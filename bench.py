"""Benchmark of the BPC-PaCo codec on one MI355X.

One step = encode + decode of the workload's images, with coefficients, codestreams
and LUTs resident in HBM.  The default workload is the plain per-image launch on a
4096x4096 image (4096 codeblocks of 64x64, 5-level DWT, 16 magnitude bitplanes,
3 coding passes), synthetic data drawn as the workload definition states.

Prints one JSON line (metric, value in Msamples/s, roofline of the dominant kernel
with its HBM traffic measured by two rocprofv3 counter passes over the same
kernels, CPU baseline from the oracle on the host cores) whose last object,
`summary`, holds the headline ratios.  The per-setting tables (autotuner sweep,
served-rate sweep) go to a detail JSON file (--detail-out).
"""
import argparse
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda_bpc-paco_amd"))

import bpcpaco as bpc  # noqa: E402
import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
DETAIL = {}  # per-setting tables, written to --detail-out
METRIC = "encode+decode Msamples/s (4096² 5-lvl DWT), batched/served vs plain; p50/p99"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(size, cp3, nbp, mag_bits, ll_bits, detail_exp, seed, levels=5):
    img = synth.synth_image(size, size, levels, seed, ll_bits=ll_bits, detail_exp=detail_exp, mag_bits=mag_bits)
    train = synth.synth_image(size, size, levels, seed + 1, ll_bits=ll_bits, detail_exp=detail_exp,
                              mag_bits=mag_bits)
    p = bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
    lut = bpc.train_lut(train, bpc.params(cpass3=cp3, dwt_levels=levels, lut_n_bitplanes=nbp))
    return img, lut, p


KERNEL_KEYS = {"encoder": "bpc_codec_kernel<0,", "decoder": "bpc_codec_kernel<1,", "copy": "copy_kernel"}
CALIBRATION_MIB = 256


def traffic_child(args):
    """Child run under rocprofv3 --pmc: two encode+decode steps of the plain
    workload (data saved by the parent) and two plain copies of 256 MiB (the
    counters' calibration).  Prints nothing on success."""
    d = np.load(args.traffic_child)
    img, lut_arr = d["img"], d["lut"]
    size = img.shape[0]
    p = bpc.params(cpass3=int(d["cp3"]), lut_n_bitplanes=int(d["nbp"]), layout=bpc.LAYOUT_TILED)
    lut = bpc.Lut.from_array(lut_arr)
    d_in = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
    d_cs = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    d_out = bpc.DeviceBuffer(size * size)
    d_ncw = bpc.DeviceBuffer(p.n_codeblocks(size, size))
    for _ in range(2):
        bpc.encode_device(d_in, d_cs, size, size, lut, p, 1, d_ncw)
        bpc.decode_device(d_cs, d_out, size, size, lut, p)
    if bpc.status_take() != 0:
        raise RuntimeError("traffic child: codec status")
    n = CALIBRATION_MIB << 20
    a, b = bpc.DeviceBuffer(n // 4), bpc.DeviceBuffer(n // 4)
    a.fill_bytes(1)
    for _ in range(2):
        b.copy_from(a)
    bpc.device_sync()


def counter_medians(outdir, counter):
    """Median counter value per kernel (encoder, decoder, copy) from a rocprofv3
    counter-collection CSV."""
    import csv
    vals = {k: [] for k in KERNEL_KEYS}
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("no counter collection output in %s" % outdir)
    for path in files:
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                for k, key in KERNEL_KEYS.items():
                    if key in row.get("Kernel_Name", ""):
                        vals[k].append(float(row["Counter_Value"]))
    if not all(vals.values()):
        raise RuntimeError("counter %s missing for %s" % (counter, [k for k, v in vals.items() if not v]))
    return {k: float(np.median(v)) for k, v in vals.items()}


def measure_traffic(img, lut_arr, cp3, nbp):
    """HBM bytes per launch of the encoder and the decoder: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two rocprofv3 --pmc runs of a child process,
    each scaled by what the counter reads for a plain copy of known size."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    tmp = tempfile.mkdtemp(prefix="bpc_traffic_")
    try:
        data = os.path.join(tmp, "plain.npz")
        np.savez(data, img=img, lut=lut_arr, cp3=cp3, nbp=nbp)
        med = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--output-format", "csv", "-d", out,
                   "-o", "run", "--", sys.executable, os.path.abspath(__file__), "--traffic-child", data]
            log("traffic: %s pass" % counter)
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp)
            if r.returncode != 0:
                return None, "rocprofv3 %s pass exited %d: %s" % (counter, r.returncode, r.stderr[-300:])
            med[counter] = counter_medians(out, counter)
        cal = float(CALIBRATION_MIB << 20)
        f_scale = cal / (med["FETCH_SIZE"]["copy"] * 1024.0)
        w_scale = cal / (med["WRITE_SIZE"]["copy"] * 1024.0)
        res = {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)}
        for k in ("encoder", "decoder"):
            res[k] = {"fetch_kib": med["FETCH_SIZE"][k], "write_kib": med["WRITE_SIZE"][k],
                      "hbm_bytes": int(round(1024.0 * (med["FETCH_SIZE"][k] * f_scale +
                                                       med["WRITE_SIZE"][k] * w_scale)))}
        return res, None
    except Exception as e:  # the line is still printed, with the reason
        return None, "traffic measurement failed: %s" % e
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(img, lut_arr, p, cp3, nbp, budget_s=20.0):
    """Oracle (C, OpenMP over codeblocks, built -O3 -march=native on this host)
    encode+decode of the same image on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        flags = O.use_native()
    except Exception as e:  # no compiler here: the portable build
        log("native oracle build failed (%s); portable build" % e)
        flags = "-O3 -march=x86-64-v2 -fopenmp"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    op = O.params(cpass3=cp3, dwt_levels=p.dwt_levels, lut_n_bitplanes=nbp)
    Y, X = img.shape
    reps, total, samples = 0, 0.0, 0
    enc_stream = None
    while total < budget_s and reps < 50:
        t0 = time.perf_counter()
        st, stream, ncw = O.encode(img, lut_arr, op, threads)
        t1 = time.perf_counter()
        st2, rec = O.decode(stream, X, Y, lut_arr, op, threads)
        t2 = time.perf_counter()
        assert st == 0 and st2 == 0
        total += t2 - t0
        samples += img.size
        reps += 1
        enc_stream = stream
    return {"value": samples / total / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "build": flags,
            "sample": "%d x encode+decode of the full %dx%d image (oracle, OpenMP over codeblocks)" % (reps, X, Y)}, \
        enc_stream


def percentiles(xs):
    return [round(float(v), 4) for v in np.percentile(xs, [50, 99])]


def bench_batched(args, check=True):
    """96 images (32 each of 2048^2, 3072^2, 4096^2), 5-level DWT, 14 magnitude
    bitplanes, a LUT per image trained on its own seeded training image, 2- and
    3-pass alternating: one persistent launch (graph replay) per direction, against
    96 plain per-image launches per direction."""
    from concurrent.futures import ThreadPoolExecutor
    sizes = [2048] * 32 + [3072] * 32 + [4096] * 32
    nbp = 15

    def setup(k):
        size, cp3 = sizes[k], k % 2
        seed = args.seed + 1000 + 2 * k
        img = synth.synth_image(size, size, 5, seed, mag_bits=14)
        train = synth.synth_image(size, size, 5, seed + 1, mag_bits=14)
        lut = bpc.train_lut(train, bpc.params(cpass3=cp3, lut_n_bitplanes=nbp), nthreads=2)
        return img, lut

    log("batched_96: generating 96 images and LUTs")
    with ThreadPoolExecutor(max_workers=8) as ex:
        data = list(ex.map(setup, range(len(sizes))))
    items = []
    for k, (img, lut_arr) in enumerate(data):
        size = sizes[k]
        p = bpc.params(cpass3=k % 2, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
        items.append(dict(size=size, p=p, lut=bpc.Lut.from_array(lut_arr),
                          d_in=bpc.DeviceBuffer.from_array(synth.to_tiled(img)),
                          d_cs=bpc.DeviceBuffer(bpc.stream_words(size, size, p)),
                          d_out=bpc.DeviceBuffer(size * size), d_ncw=bpc.DeviceBuffer(p.n_codeblocks(size, size))))
    total_cbs = sum(it["p"].n_codeblocks(it["size"], it["size"]) for it in items)
    samples = sum(it["size"] ** 2 for it in items)
    stream = bpc.Stream()
    d_status = bpc.DeviceBuffer(1, np.uint32)
    d_status.fill_bytes(0)
    bpc.device_sync()
    batch = bpc.Batch(len(items), total_cbs)
    enc = [bpc.make_image(it["d_in"], it["d_cs"], it["lut"], it["size"], it["size"], it["p"], it["d_ncw"])
           for it in items]
    dec = [bpc.make_image(it["d_cs"], it["d_out"], it["lut"], it["size"], it["size"], it["p"]) for it in items]
    ev = [bpc.Event() for _ in range(3)]

    def use(t):
        batch.set_tuning(t["codeblocks_per_workgroup"] // 2, 0 if t["waves_per_simd"] == "max" else t["waves_per_simd"])
        batch.set_vgpr_cap(t["vgpr_cap_waves_per_simd"])

    chosen = {}

    def batched_step(graph=True):
        ev[0].record(stream)
        if chosen:
            use(chosen["encode"])
        batch.run(bpc.CODE, enc, use_graph=graph, d_status=d_status, stream=stream)
        ev[1].record(stream)
        if chosen:
            use(chosen["decode"])
        batch.run(bpc.DECODE, dec, use_graph=graph, d_status=d_status, stream=stream)
        ev[2].record(stream)
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])

    # autotuner, per direction: codeblocks per workgroup (2 per wave) x resident
    # waves per SIMD x register budget (compiler's, or capped for 5..8 waves per SIMD)
    tune = []
    grid = [(wpb, wps, cap) for wpb in (1, 2, 4) for wps in (0, 2, 3) for cap in (0, 5, 6, 7, 8)]
    if args.no_autotune:  # the setting the sweep picks on MI355X, for short (profiled) runs
        grid = [(4, 0, 8)]
    for wpb, wps, cap in grid:
        t = {"codeblocks_per_workgroup": 2 * wpb, "waves_per_simd": wps or "max", "vgpr_cap_waves_per_simd": cap}
        use(t)
        batched_step(graph=False)
        e, d = batched_step(graph=False)
        t.update({"encode_ms": round(e, 3), "decode_ms": round(d, 3)})
        tune.append(t)
    keys = ("codeblocks_per_workgroup", "waves_per_simd", "vgpr_cap_waves_per_simd")
    chosen["encode"] = {k: min(tune, key=lambda t: t["encode_ms"])[k] for k in keys}
    chosen["decode"] = {k: min(tune, key=lambda t: t["decode_ms"])[k] for k in keys}
    best = dict(chosen)
    rc = bpc.status_take(d_status, stream)
    if rc != 0:
        raise RuntimeError("batched codec status %d: %s" % (rc, bpc.lib().bpc_last_error().decode()))
    if check:
        log("batched_96: checking round trips")
        for k, it in enumerate(items):
            rec = it["d_out"].download(stream=stream)
            if not np.array_equal(synth.from_tiled(rec, it["size"], it["size"]), data[k][0]):
                raise RuntimeError("batched round trip failed on image %d" % k)
    del data
    steps = max(2, args.steps // 4)
    for _ in range(args.warmup):
        batched_step()
    lat = []
    bpc.device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e, d = batched_step()
        lat.append(e + d)
    bpc.device_sync()
    wall_b = time.perf_counter() - t0
    # plain per-image launches, same images
    def plain_step():
        ev[0].record(stream)
        for it in items:
            bpc.encode_device(it["d_in"], it["d_cs"], it["size"], it["size"], it["lut"], it["p"], 1, it["d_ncw"],
                              d_status, stream)
        ev[1].record(stream)
        for it in items:
            bpc.decode_device(it["d_cs"], it["d_out"], it["size"], it["size"], it["lut"], it["p"], d_status, stream)
        ev[2].record(stream)
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])
    for _ in range(args.warmup):
        plain_step()
    plat = []
    bpc.device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e, d = plain_step()
        plat.append(e + d)
    bpc.device_sync()
    wall_p = time.perf_counter() - t0
    rc = bpc.status_take(d_status, stream)
    if rc != 0:
        raise RuntimeError("codec status %d" % rc)
    vb = samples * steps / wall_b / 1e6
    vp = samples * steps / wall_p / 1e6
    out = {"workload": "batched_96 (32x2048^2 + 32x3072^2 + 32x4096^2, 2/3-pass alternating, LUT per image)",
           "images": len(items), "codeblocks": total_cbs, "samples": samples, "steps": steps,
           "batched_msamples_s": round(vb, 2), "plain_96_launches_msamples_s": round(vp, 2),
           "batched_vs_plain": round(vb / vp, 3),
           "batched_step_ms_p50_p99": percentiles(lat), "plain_step_ms_p50_p99": percentiles(plat),
           "autotune_settings": len(tune), "chosen": best,
           "chosen_encode_ms": min(t["encode_ms"] for t in tune), "chosen_decode_ms": min(t["decode_ms"] for t in tune),
           "graph_replay": True}
    DETAIL["batched_autotune"] = tune
    batch.destroy()
    for it in items:
        for k in ("d_in", "d_cs", "d_out", "d_ncw"):
            it[k].free()
        it["lut"].free()
    return out


def bench_served(args):
    """Open-loop Poisson stream of 50% encode / 50% decode requests, sizes uniform
    over {2048^2, 3072^2, 4096^2, 6144^2}, 14-16 bitplanes, 2 or 3 passes; decode
    requests replay codestreams produced earlier in the run and kept in HBM.  The
    continuous-batching server against a plain one-request-at-a-time loop."""
    import random
    import threading
    sizes = [2048, 3072, 4096, 6144]
    variants = [(0, 15, 14), (1, 16, 16)]  # (cpass3, LUT bitplanes, magnitude bitplanes)
    pool = []
    log("served: generating request pool")
    for si, size in enumerate(sizes):
        for vi, (cp3, nbp, mag) in enumerate(variants):
            seed = args.seed + 5000 + 10 * si + 2 * vi
            img = synth.synth_image(size, size, 5, seed, mag_bits=mag)
            lut_arr = bpc.train_lut(synth.synth_image(size, size, 5, seed + 1, mag_bits=mag),
                                    bpc.params(cpass3=cp3, lut_n_bitplanes=nbp))
            p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp, layout=bpc.LAYOUT_TILED)
            lut = bpc.Lut.from_array(lut_arr)
            d_img = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
            d_cs = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
            bpc.encode_device(d_img, d_cs, size, size, lut, p)  # the codestream decode requests replay
            pool.append(dict(size=size, p=p, lut=lut, d_img=d_img, d_cs=d_cs))
    if bpc.status_take() != 0:
        raise RuntimeError("served pool encode failed")
    d_tmp_cs = bpc.DeviceBuffer(max(bpc.stream_words(it["size"], it["size"], it["p"]) for it in pool))
    d_tmp_img = bpc.DeviceBuffer(max(it["size"] ** 2 for it in pool))
    stream = bpc.Stream()

    def make_plan(n, seed):
        r = random.Random(seed)
        return [(r.randrange(len(pool)), r.random() < 0.5) for _ in range(n)]

    # plain loop: one request at a time, per-image launches
    def plain_loop(plan):
        t0 = time.perf_counter()
        lat = []
        for j, enc in plan:
            it = pool[j]
            t = time.perf_counter()
            if enc:
                bpc.encode_device(it["d_img"], d_tmp_cs, it["size"], it["size"], it["lut"], it["p"], stream=stream)
            else:
                bpc.decode_device(it["d_cs"], d_tmp_img, it["size"], it["size"], it["lut"], it["p"], stream=stream)
            stream.sync()
            lat.append((time.perf_counter() - t) * 1e3)
        wall = time.perf_counter() - t0
        return sum(pool[j]["size"] ** 2 for j, _ in plan) / wall / 1e6, lat

    plain_loop(make_plan(8, 1))
    n_req = args.served_requests
    plain_rate, plain_lat = plain_loop(make_plan(n_req, 2))
    mean_samples = float(np.mean([it["size"] ** 2 for it in pool]))
    capacity_rps = plain_rate * 1e6 / mean_samples  # one-at-a-time service rate
    srv = bpc.Server(max_inflight=3, max_batch_images=64)
    sweep = []
    # offered rate as a multiple of the plain loop's capacity: coarse below 2x, fine
    # between 2x and 3.5x where the server's knee lies
    mults = (0.5, 1.0, 1.5, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5, 4.0)
    try:
        for mult in mults:
            rate = mult * capacity_rps
            plan = make_plan(n_req, int(mult * 100))
            rng = random.Random(int(mult * 1000))
            gaps = [rng.expovariate(rate) for _ in plan]
            ids = [None] * len(plan)
            done_t = [None] * len(plan)
            results = [None] * len(plan)

            def waiter():
                for k in range(len(plan)):
                    while ids[k] is None:
                        time.sleep(0.0002)
                    results[k] = srv.wait(ids[k], timeout_ms=120000)
                    done_t[k] = time.perf_counter()
                    srv.release(ids[k])

            th = threading.Thread(target=waiter)
            th.start()
            t0 = time.perf_counter()
            t_next = t0
            try:
                for k, (j, enc) in enumerate(plan):
                    t_next += gaps[k]
                    dt = t_next - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    it = pool[j]
                    ids[k] = srv.submit(bpc.CODE if enc else bpc.DECODE, it["d_img"] if enc else it["d_cs"],
                                        it["lut"], it["size"], it["size"], it["p"])
            finally:
                th.join()
            if any(r.status != 0 for r in results):
                raise RuntimeError("served request failed")
            t_last_submit = t_next
            wall = max(done_t) - t0
            lat = [r.latency_us / 1e3 for r in results]
            samples = sum(pool[j]["size"] ** 2 for j, _ in plan)
            offered = samples / (t_last_submit - t0) / 1e6
            sweep.append({"offered_x_plain": mult, "offered_rps": round(rate, 1),
                          "offered_msamples_s": round(offered, 1),
                          "sustained_msamples_s": round(samples / wall / 1e6, 1),
                          "latency_ms_p50_p99": percentiles(lat), "requests": len(plan)})
            log("served x%.2f: offered %.0f, sustained %.0f Msamples/s, p99 %.1f ms" % (
                mult, offered, sweep[-1]["sustained_msamples_s"], sweep[-1]["latency_ms_p50_p99"][1]))
        st = srv.stats()
        batches = int(st.batches)
    finally:
        srv.destroy()
    DETAIL["served_sweep"] = sweep
    # knee: the highest offered rate the server keeps up with (sustained >= 95% of offered)
    kept = [r for r in sweep if r["sustained_msamples_s"] >= 0.95 * r["offered_msamples_s"]]
    knee = max(kept, key=lambda r: r["offered_msamples_s"]) if kept else min(sweep, key=lambda r: r["offered_rps"])
    best = max(sweep, key=lambda r: r["sustained_msamples_s"])
    out = {"workload": "served (Poisson, 50% encode / 50% decode, sizes {2048,3072,4096,6144}^2, 14-16 bitplanes, "
                       "2/3 passes)",
           "plain_loop_msamples_s": round(plain_rate, 1), "plain_loop_latency_ms_p50_p99": percentiles(plain_lat),
           "requests_per_rate": n_req, "rates": len(sweep),
           "knee": {"offered_msamples_s": knee["offered_msamples_s"],
                    "sustained_msamples_s": knee["sustained_msamples_s"],
                    "latency_ms_p50_p99": knee["latency_ms_p50_p99"],
                    "vs_plain_loop": round(knee["sustained_msamples_s"] / plain_rate, 3)},
           "max_sustained": {"offered_msamples_s": best["offered_msamples_s"],
                             "sustained_msamples_s": best["sustained_msamples_s"],
                             "latency_ms_p50_p99": best["latency_ms_p50_p99"],
                             "vs_plain_loop": round(best["sustained_msamples_s"] / plain_rate, 3)},
           "batches_launched": batches,
           "server": {"max_inflight": 3, "max_batch_images": 64, "admission": "device free memory less 4 GiB"}}
    for it in pool:
        it["d_img"].free()
        it["d_cs"].free()
        it["lut"].free()
    d_tmp_cs.free()
    d_tmp_img.free()
    return out


def bench_large(args, check=True):
    """16384x16384 (65536 codeblocks, 1 GiB int32), 20 magnitude bitplanes (LL
    uniform on [0, 2^19), detail scale 2^(14-d)), 3 passes, 8192-word regions,
    batched persistent path; oracle on a seeded 1% codeblock sample."""
    size, nbp, stride = 16384, 20, 8192
    log("large_16384: generating image and LUT")
    img = synth.synth_image(size, size, 5, args.seed + 9000, ll_bits=19, detail_exp=14, mag_bits=nbp)
    train = synth.synth_image(size // 4, size // 4, 5, args.seed + 9001, ll_bits=17, detail_exp=12, mag_bits=nbp)
    lut_arr = bpc.train_lut(train, bpc.params(cpass3=1, lut_n_bitplanes=nbp))
    del train
    p = bpc.params(cpass3=1, lut_n_bitplanes=nbp, cb_stride_words=stride, layout=bpc.LAYOUT_TILED)
    ncb = p.n_codeblocks(size, size)
    lut = bpc.Lut.from_array(lut_arr)
    d_in = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
    d_cs = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    d_out = bpc.DeviceBuffer(size * size)
    d_ncw = bpc.DeviceBuffer(ncb)
    d_status = bpc.DeviceBuffer(1, np.uint32)
    d_status.fill_bytes(0)
    stream = bpc.Stream()
    bpc.device_sync()
    batch = bpc.Batch(1, ncb)
    batch.set_vgpr_cap(args.vgpr_cap)
    enc = [bpc.make_image(d_in, d_cs, lut, size, size, p, d_ncw)]
    dec = [bpc.make_image(d_cs, d_out, lut, size, size, p)]
    ev = [bpc.Event() for _ in range(3)]

    def step():
        ev[0].record(stream)
        batch.run(bpc.CODE, enc, use_graph=True, d_status=d_status, stream=stream)
        ev[1].record(stream)
        batch.run(bpc.DECODE, dec, use_graph=True, d_status=d_status, stream=stream)
        ev[2].record(stream)
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])

    for _ in range(max(1, args.warmup)):
        step()
    if bpc.status_take(d_status, stream) != 0:
        raise RuntimeError("large codec status: %s" % bpc.lib().bpc_last_error().decode())
    exact = None
    if check:
        rec = synth.from_tiled(d_out.download(stream=stream), size, size)
        if not np.array_equal(rec, img):
            raise RuntimeError("large round trip failed")
        del rec
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cs = d_cs.download(stream=stream).reshape(ncb, stride)
        sample = np.sort(np.random.default_rng(args.seed).choice(ncb, size=ncb // 100, replace=False)).astype(np.int32)
        st, ostream, _ = O.encode_list(img, sample, lut_arr, O.params(cpass3=1, lut_n_bitplanes=nbp,
                                                                      cb_stride_words=stride))
        ost = ostream.reshape(ncb, stride)
        exact = bool(st == 0 and all(np.array_equal(cs[c], ost[c]) for c in sample))
        del cs, ostream, ost
        if not exact:
            raise RuntimeError("large sample differs from the oracle")
    steps = max(2, args.steps // 4)
    times = []
    bpc.device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e, d = step()
        times.append((e, d))
    bpc.device_sync()
    wall = time.perf_counter() - t0
    ncw = d_ncw.download(stream=stream)
    out = {"workload": "large_16384 (65536 codeblocks, 20 bitplanes, 3-pass, 8192-word regions, batched path)",
           "vgpr_cap_waves_per_simd": args.vgpr_cap,
           "msamples_s": round(size * size * steps / wall / 1e6, 1),
           "encode_ms": round(float(np.mean([e for e, _ in times])), 3),
           "decode_ms": round(float(np.mean([d for _, d in times])), 3),
           "bits_per_sample": round(float(ncw.sum()) * 16 / (size * size), 3),
           "max_codewords_per_codeblock": int(ncw.max()),
           "oracle_sample_codeblocks": ncb // 100 if check else 0, "bit_exact_sample_and_round_trip": exact}
    batch.destroy()
    for b in (d_in, d_cs, d_out, d_ncw, d_status):
        b.free()
    lut.free()
    return out


def bench_container(args, d_cs, d_ncw, ncb, stride, stream, plain):
    """Compact container of the plain workload's codestream: index + pack (regions ->
    MSB bytes, offsets, uint16 codewords) and unpack (back to -1-filled regions), with
    the round trip checked.  HBM-bound; algorithmic bytes per codeblock: pack reads
    the MSB, count and used words (4 * (2 + n)) and writes 1 + 8 + 2n; unpack reads
    1 + 8 + 2n and writes the whole region (4 * stride).  Then the decoder reading the
    compact form directly against the region decoder (and unpack + decode), both
    checked against the original coefficients.  plain: (lut, params, size, d_in)."""
    msb = bpc.DeviceBuffer(ncb, np.uint8)
    off = bpc.DeviceBuffer(ncb + 1, np.uint64)
    bpc.compact_index(d_cs, ncb, stride, msb, off, d_ncw=d_ncw, stream=stream)
    total = int(bpc.download_ptr(off.ptr.value + 8 * ncb, 1, np.uint64, stream)[0])
    payload = bpc.DeviceBuffer(max(total, 1), np.uint16)
    back = bpc.DeviceBuffer(ncb * stride)
    c = bpc.Compact(msb, off, payload, ncb, total)
    ev = [bpc.Event() for _ in range(3)]
    tp, tu = [], []
    for k in range(max(1, args.warmup) + max(5, args.steps)):
        ev[0].record(stream)
        bpc.compact_index(d_cs, ncb, stride, msb, off, d_ncw=d_ncw, stream=stream)
        bpc.compact_pack_into(d_cs, ncb, stride, off, payload, total, stream=stream)
        ev[1].record(stream)
        bpc.compact_unpack(c, back, stride, stream=stream)
        ev[2].record(stream)
        if k >= max(1, args.warmup):
            tp.append(ev[0].elapsed_ms(ev[1]))
            tu.append(ev[1].elapsed_ms(ev[2]))
    if bpc.status_take(None, stream) != 0:
        raise RuntimeError("container status: %s" % bpc.lib().bpc_last_error().decode())
    exact = bool(np.array_equal(back.download(stream=stream), d_cs.download(stream=stream)))
    if not exact:
        raise RuntimeError("compact round trip differs")
    pack_ms, unpack_ms = float(np.median(tp)), float(np.median(tu))
    pack_bytes = 4.0 * (2 * ncb + total) + 9.0 * ncb + 2.0 * total
    unpack_bytes = 9.0 * ncb + 2.0 * total + 4.0 * stride * ncb
    # decoding: from the regions, and straight from the compact form
    lut, p, size, d_in = plain
    rec_r, rec_c = bpc.DeviceBuffer(size * size), bpc.DeviceBuffer(size * size)
    tr, tc = [], []
    for k in range(max(1, args.warmup) + max(5, args.steps)):
        ev[0].record(stream)
        bpc.decode_device(d_cs, rec_r, size, size, lut, p, stream=stream)
        ev[1].record(stream)
        bpc.decode_compact_device(c, rec_c, size, size, lut, p, stream=stream)
        ev[2].record(stream)
        if k >= max(1, args.warmup):
            tr.append(ev[0].elapsed_ms(ev[1]))
            tc.append(ev[1].elapsed_ms(ev[2]))
    if bpc.status_take(None, stream) != 0:
        raise RuntimeError("decode status: %s" % bpc.lib().bpc_last_error().decode())
    orig = d_in.download(stream=stream)
    dec_exact = bool(np.array_equal(rec_c.download(stream=stream), orig) and
                     np.array_equal(rec_r.download(stream=stream), orig))
    if not dec_exact:
        raise RuntimeError("decode from the compact form differs from the original")
    dec_r_ms, dec_c_ms = float(np.median(tr)), float(np.median(tc))
    for b in (payload, back, msb, off, rec_r, rec_c):
        b.free()
    return {"workload": "compact container of the plain_4096_3pass codestream",
            "compact_bytes": c.nbytes, "reference_layout_bytes": 4 * stride * ncb,
            "compact_fraction": round(c.nbytes / (4.0 * stride * ncb), 4),
            "pack_ms": round(pack_ms, 4), "unpack_ms": round(unpack_ms, 4),
            "pack_gbs": round(pack_bytes / pack_ms / 1e6, 1), "unpack_gbs": round(unpack_bytes / unpack_ms / 1e6, 1),
            "unpack_hbm_frac": round(unpack_bytes / unpack_ms / 1e6 / HBM_PEAK_GBS, 4),
            "round_trip_exact": exact,
            "decode_regions_ms": round(dec_r_ms, 4), "decode_compact_ms": round(dec_c_ms, 4),
            "unpack_then_decode_ms": round(unpack_ms + dec_r_ms, 4), "decode_compact_exact": dec_exact}


def bench_dwt(args, check=True):
    """Reversible 5/3 transform of a 4096x4096 int32 image, 5 levels, forward and
    inverse on the device.  HBM-bound; algorithmic bytes: every level reads and
    writes its band once, 8 * X * Y * (1 + 1/4 + ... + 1/4^4) per direction."""
    size, levels = 4096, 5
    rng = np.random.default_rng(args.seed + 500)
    img = rng.integers(0, 1 << 12, (size, size)).astype(np.int32)
    stream = bpc.Stream()
    d_img = bpc.DeviceBuffer.from_array(img)
    d_coef = bpc.DeviceBuffer(size * size)
    d_rec = bpc.DeviceBuffer(size * size)
    bpc.device_sync()
    ev = [bpc.Event() for _ in range(3)]
    tf, ti = [], []
    for k in range(max(1, args.warmup) + max(5, args.steps)):
        ev[0].record(stream)
        bpc.dwt53_forward_device(d_img, d_coef, size, size, levels, stream)
        ev[1].record(stream)
        bpc.dwt53_inverse_device(d_coef, d_rec, size, size, levels, stream)
        ev[2].record(stream)
        if k >= max(1, args.warmup):
            tf.append(ev[0].elapsed_ms(ev[1]))
            ti.append(ev[1].elapsed_ms(ev[2]))
    rec = d_rec.download(stream=stream).reshape(size, size)
    if not np.array_equal(rec, img):
        raise RuntimeError("inverse(forward(x)) != x")
    exact = None
    if check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import dwt53
        exact = bool(np.array_equal(d_coef.download(stream=stream).reshape(size, size), dwt53.forward(img, levels)))
        if not exact:
            raise RuntimeError("device transform differs from the oracle")
    for b in (d_img, d_coef, d_rec):
        b.free()
    fwd_ms, inv_ms = float(np.median(tf)), float(np.median(ti))
    alg = 8.0 * size * size * sum(4.0 ** -l for l in range(levels))
    return {"workload": "dwt53_4096 (4096x4096 int32, 5 levels)",
            "forward_ms": round(fwd_ms, 4), "inverse_ms": round(inv_ms, 4),
            "forward_gbs": round(alg / fwd_ms / 1e6, 1), "inverse_gbs": round(alg / inv_ms / 1e6, 1),
            "forward_hbm_frac": round(alg / fwd_ms / 1e6 / HBM_PEAK_GBS, 4),
            "msamples_s_forward_inverse": round(size * size / (fwd_ms + inv_ms) / 1e3, 1),
            "round_trip_exact": True, "bit_exact_vs_oracle": exact}


def pinned_int32(n):
    """A page-locked host array of n int32 (hipHostMalloc, as the reference example's
    cudaMallocHost, main_example.cu:508), or None when the HIP runtime is not found."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return None, None
    ptr = ctypes.c_void_p()
    if hip.hipHostMalloc(ctypes.byref(ptr), ctypes.c_size_t(4 * n), ctypes.c_uint(0)) != 0 or not ptr.value:
        return None, None
    arr = np.ctypeslib.as_array((ctypes.c_int32 * n).from_address(ptr.value))

    def free():
        hip.hipHostFree(ptr)
    return arr, free


def bench_drop_in(args):
    """The drop-in CUDA_BPC path end to end (BPC_PaCo.cu:1305-1325): host buffer in,
    device buffers allocated, H2D copy, kernel, D2H copy, freed, per call; the plain
    4096^2 3-pass image in raster layout, from pageable and from pinned host memory."""
    size, cp3, nbp = 4096, 1, 16
    log("drop_in: generating image and training LUT")
    img, lut_arr, _ = make_workload(size, cp3, nbp, mag_bits=16, ll_bits=13, detail_exp=9, seed=args.seed)
    p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp)
    lut = bpc.Lut.from_array(lut_arr)
    n = max(size * size, bpc.stream_words(size, size, p))
    out = {"workload": "drop_in_4096_3pass (bpc_run: allocation, copies and kernel per call, raster layout)"}
    pinned, free = pinned_int32(n)
    kinds = [("pageable", np.zeros(n, np.int32))] + ([("pinned", pinned)] if pinned is not None else [])
    try:
        for kind, buf in kinds:
            times = []
            for rep in range(6):
                buf[: size * size] = img.ravel()
                t0 = time.perf_counter()
                bpc.run(bpc.CODE, buf, size, size, lut, p)
                t1 = time.perf_counter()
                bpc.run(bpc.DECODE, buf, size, size, lut, p)
                t2 = time.perf_counter()
                if not np.array_equal(buf[: size * size].reshape(size, size), img):
                    raise RuntimeError("drop-in round trip differs (%s)" % kind)
                if rep:
                    times.append((t1 - t0, t2 - t1))
            e, d = np.median(times, axis=0)
            out[kind] = {"encode_ms": round(e * 1e3, 3), "decode_ms": round(d * 1e3, 3),
                         "msamples_s": round(size * size / (e + d) / 1e6, 1), "round_trip_exact": True}
    finally:
        del kinds
        if free is not None:
            del pinned
            free()
        lut.free()
    return out


def bench_int16(args):
    """int16 coefficient path: a 4096x4096 image of at most 15 magnitude bits, 3
    passes, coded from int16 and from int32 coefficients of equal value (the
    codestreams must be identical), encode + decode each."""
    size, cp3, nbp = 4096, 1, 16
    img, lut_arr, p = make_workload(size, cp3, nbp, mag_bits=15, ll_bits=13, detail_exp=9, seed=args.seed + 700)
    lut = bpc.Lut.from_array(lut_arr)
    tiled = synth.to_tiled(img)
    stream = bpc.Stream()
    d16 = bpc.DeviceBuffer.from_array(tiled.astype(np.int16))
    d32 = bpc.DeviceBuffer.from_array(tiled)
    cs16 = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    cs32 = bpc.DeviceBuffer(bpc.stream_words(size, size, p))
    o16 = bpc.DeviceBuffer(size * size, np.int16)
    o32 = bpc.DeviceBuffer(size * size)
    bpc.device_sync()
    ev = [bpc.Event() for _ in range(3)]

    def run(i16):
        ev[0].record(stream)
        if i16:
            bpc.encode_device_i16(d16, cs16, size, size, lut, p, stream=stream)
        else:
            bpc.encode_device(d32, cs32, size, size, lut, p, stream=stream)
        ev[1].record(stream)
        if i16:
            bpc.decode_device_i16(cs16, o16, size, size, lut, p, stream=stream)
        else:
            bpc.decode_device(cs32, o32, size, size, lut, p, stream=stream)
        ev[2].record(stream)
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])

    res = {}
    for i16 in (True, False):
        t = [run(i16) for _ in range(max(1, args.warmup) + max(5, args.steps // 2))][max(1, args.warmup):]
        e, d = float(np.median([x for x, _ in t])), float(np.median([y for _, y in t]))
        res["int16" if i16 else "int32"] = {"encode_ms": round(e, 4), "decode_ms": round(d, 4),
                                            "msamples_s": round(size * size / (e + d) / 1e3, 1)}
    if bpc.status_take(None, stream) != 0:
        raise RuntimeError("int16 leg status: %s" % bpc.lib().bpc_last_error().decode())
    same = bool(np.array_equal(cs16.download(stream=stream), cs32.download(stream=stream)))
    exact = bool(np.array_equal(o16.download(stream=stream), tiled.astype(np.int16)))
    if not (same and exact):
        raise RuntimeError("int16 path differs from the int32 path")
    for b in (d16, d32, cs16, cs32, o16, o32):
        b.free()
    lut.free()
    res.update({"workload": "plain_4096_3pass_15bit (int16 vs int32 coefficients, same data)",
                "coefficient_bytes": {"int16": 2 * size * size, "int32": 4 * size * size},
                "same_codestream": same, "round_trip_exact": exact})
    return res


def bench_train(args):
    """LUT training on the device (symbol statistics of the encoder's modelling, no
    coding) for the plain workload's training image: 4096x4096, 3 passes, 16 LUT
    bitplanes; the host trainer on all host threads beside it, same table required."""
    size, cp3, nbp = 4096, 1, 16
    log("train_4096_3pass: generating the training image")
    train = synth.synth_image(size, size, 5, args.seed + 1, ll_bits=13, detail_exp=9, mag_bits=16)
    p = bpc.params(cpass3=cp3, lut_n_bitplanes=nbp)
    n = bpc.lut_len(p)
    stream = bpc.Stream()
    d_img = bpc.DeviceBuffer.from_array(train)
    d_lut = bpc.DeviceBuffer(n)
    bpc.device_sync()
    ev = [bpc.Event() for _ in range(2)]
    times = []
    for k in range(max(1, args.warmup) + max(3, args.steps // 2)):
        ev[0].record(stream)
        bpc.train_lut_device(d_img, size, size, p, d_lut=d_lut, stream=stream)
        ev[1].record(stream)
        ms = ev[0].elapsed_ms(ev[1])
        if k >= max(1, args.warmup):
            times.append(ms)
    dev_lut = d_lut.download(stream=stream)
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    host_lut = bpc.train_lut(train, p, nthreads=threads)
    host_s = time.perf_counter() - t0
    exact = bool(np.array_equal(dev_lut, host_lut))
    if not exact:
        raise RuntimeError("device LUT differs from the host trainer's")
    gpu_ms = float(np.median(times))
    d_img.free()
    d_lut.free()
    return {"workload": "train_4096_3pass (4096x4096 training image, 16 LUT bitplanes, %d entries)" % n,
            "gpu_ms": round(gpu_ms, 4), "gpu_msamples_s": round(size * size / gpu_ms / 1e3, 1),
            "host_trainer_ms": round(host_s * 1e3, 2), "host_threads": threads,
            "host_msamples_s": round(size * size / host_s / 1e6, 1), "identical_to_host_trainer": exact}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle baseline and parity check")
    ap.add_argument("--workloads", default="plain,batched,served,large,train,container,dwt,int16,drop_in",
                    help="comma list of plain, batched, served, large, train, container, dwt, int16, drop_in "
                         "(plain always runs: it gives `value`)")
    ap.add_argument("--vgpr-cap", type=int, default=8,
                    help="register cap of the batched kernel for the large workload (0 or 5..8; the batched "
                         "workload's autotuner overrides it when both run)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 5 times (HBM counter calibration)")
    ap.add_argument("--served-requests", type=int, default=2000,
                    help="requests per offered rate (the drain of the last batch is a smaller share of a longer run)")
    ap.add_argument("--detail-out", default=os.path.join(ROOT, "gpurun_out", "bench_detail.json"),
                    help="JSON file for the autotuner and served-rate tables ('' = none)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched workload: one setting (8 codeblocks per workgroup, all waves, register cap 8) "
                         "instead of the 45-setting sweep")
    ap.add_argument("--traffic-child", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: every codeblock is independent and a multi-GPU run "
              "would be independent replicas; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if bpc.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        a, b = bpc.DeviceBuffer(n // 4), bpc.DeviceBuffer(n // 4)
        a.fill_bytes(1)
        for _ in range(5):
            b.copy_from(a)
        bpc.device_sync()
        print(json.dumps({"calibration_copy_bytes": n, "launches": 5}))
        return

    size, cp3, nbp = 4096, 1, 16
    log("workload plain_4096_3pass: generating image and training LUT")
    img, lut_arr, p = make_workload(size, cp3, nbp, mag_bits=16, ll_bits=13, detail_exp=9, seed=args.seed)
    X = Y = size
    ncb = p.n_codeblocks(X, Y)
    lut = bpc.Lut.from_array(lut_arr)
    stream = bpc.Stream()
    d_in = bpc.DeviceBuffer.from_array(synth.to_tiled(img))
    nw = bpc.stream_words(X, Y, p)
    d_cs = bpc.DeviceBuffer(nw)
    d_out = bpc.DeviceBuffer(X * Y)
    d_ncw = bpc.DeviceBuffer(ncb)
    d_status = bpc.DeviceBuffer(1, np.uint32)
    d_status.fill_bytes(0)
    bpc.device_sync()  # setup copies and the memset ran on the null stream
    ev = [bpc.Event() for _ in range(4)]

    def step(timed_pairs=None):
        # events only in the per-kernel pass: the wall-clock steps carry no extra packets
        if timed_pairs is not None:
            ev[0].record(stream)
        bpc.encode_device(d_in, d_cs, X, Y, lut, p, 1, d_ncw, d_status, stream)
        if timed_pairs is not None:
            ev[1].record(stream)
            ev[2].record(stream)
        bpc.decode_device(d_cs, d_out, X, Y, lut, p, d_status, stream)
        if timed_pairs is not None:
            ev[3].record(stream)
            timed_pairs.append((ev[0].elapsed_ms(ev[1]), ev[2].elapsed_ms(ev[3])))

    log("warmup %d steps" % args.warmup)
    for _ in range(args.warmup):
        step()
    rc = bpc.status_take(d_status, stream)
    if rc != 0:
        raise RuntimeError("codec status %d: %s" % (rc, bpc.lib().bpc_last_error().decode()))
    # correctness of this run: round trip, and the codestream against the oracle below
    rec = synth.from_tiled(d_out.download(stream=stream), X, Y)
    if not np.array_equal(rec, img):
        raise RuntimeError("decode(encode(x)) != x")
    ncw = d_ncw.download(stream=stream)
    gstream = d_cs.download(stream=stream)

    # timed region: K steps between two full synchronisations
    bpc.device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    stream.sync()
    bpc.device_sync()
    wall = time.perf_counter() - t0
    # per-kernel durations from HIP events on the launch stream (separate pass, same work)
    pairs = []
    for _ in range(args.steps):
        step(pairs)
    rc = bpc.status_take(d_status, stream)
    if rc != 0:
        raise RuntimeError("codec status %d" % rc)
    enc_ms = float(np.mean([a for a, _ in pairs]))
    dec_ms = float(np.mean([b for _, b in pairs]))
    enc_p50, enc_p99 = np.percentile([a for a, _ in pairs], [50, 99])
    dec_p50, dec_p99 = np.percentile([b for _, b in pairs], [50, 99])

    samples = X * Y
    ms_per_step = wall * 1e3 / args.steps
    value = samples * args.steps / wall / 1e6
    # algorithmic bytes per direction: 16384 B of coefficients + 4*(1 + n_cw) B of
    # codestream words (MSB + used slots) per codeblock
    alg_bytes = float(ncb * 16384 + 4 * (ncb + int(ncw.sum())))
    dom_name, dom_ms = ("decoder", dec_ms) if dec_ms >= enc_ms else ("encoder", enc_ms)
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    traffic, traffic_err = (None, "skipped (--no-traffic)") if args.no_traffic else \
        measure_traffic(img, lut_arr, cp3, nbp)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Laplacian/uniform DWT coefficients, LUT trained on a second seeded image)",
        "config": {"workload": "plain_4096_3pass", "mode": "plain per-image launch", "image": "4096x4096",
                   "codeblocks": ncb, "codeblock": "64x64", "dwt_levels": 5, "cpass3": cp3,
                   "magnitude_bitplanes": 16, "lut_n_bitplanes": nbp, "layout": "codeblock-contiguous",
                   "seed": args.seed},
        "encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
        "encode_ms_p50_p99": [round(float(enc_p50), 4), round(float(enc_p99), 4)],
        "decode_ms_p50_p99": [round(float(dec_p50), 4), round(float(dec_p99), 4)],
        "bits_per_sample": round(float(ncw.sum()) * 16 / samples, 4),
        "roofline": {"bound": "hbm", "kernel": "bpc_codec_kernel (%s)" % dom_name,
                     "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": traffic[dom_name]["hbm_bytes"] if traffic else None,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "traffic_by_kernel": ({k: {"hbm_bytes": traffic[k]["hbm_bytes"],
                                                "vs_algorithmic": round(traffic[k]["hbm_bytes"] / alg_bytes, 3)}
                                            for k in ("encoder", "decoder")} if traffic else None),
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a %d MiB copy "
                                        "(x%.3f / x%.3f)" % (CALIBRATION_MIB, traffic["fetch_scale"],
                                                             traffic["write_scale"]) if traffic else traffic_err)},
        "cpu_baseline": None,
        "bit_exact_vs_oracle": None,
    }
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cb, ostream = cpu_baseline(img, lut_arr, p, cp3, nbp)
        out["cpu_baseline"] = cb
        out["bit_exact_vs_oracle"] = bool(np.array_equal(gstream, ostream))
        if not out["bit_exact_vs_oracle"]:
            raise RuntimeError("GPU codestream differs from the oracle")
    wl = args.workloads.split(",")
    if "container" in wl:
        out["container"] = bench_container(args, d_cs, d_ncw, ncb, p.stride(), stream, (lut, p, size, d_in))
    for b in (d_in, d_cs, d_out, d_ncw, d_status):
        b.free()
    lut.free()
    if "batched" in wl:
        out["batched"] = bench_batched(args, check=not args.no_cpu)
        args.vgpr_cap = out["batched"]["chosen"]["encode"]["vgpr_cap_waves_per_simd"]
    if "served" in wl:
        out["served"] = bench_served(args)
    if "large" in wl:
        out["large"] = bench_large(args, check=not args.no_cpu)
    if "train" in wl:
        out["lut_trainer"] = bench_train(args)
    if "int16" in wl:
        out["int16"] = bench_int16(args)
    if "drop_in" in wl:
        out["drop_in"] = bench_drop_in(args)
    if "dwt" in wl:
        out["dwt53"] = bench_dwt(args, check=not args.no_cpu)
    # headline ratios last, so that the end of the line carries them
    summary = {"plain_msamples_s": out["value"], "roofline_frac": out["roofline"]["frac"]}
    if out["roofline"]["traffic_by_kernel"]:
        summary["traffic_vs_algorithmic"] = {k: v["vs_algorithmic"]
                                             for k, v in out["roofline"]["traffic_by_kernel"].items()}
    if out.get("cpu_baseline"):
        summary["cpu_baseline_msamples_s"] = round(out["cpu_baseline"]["value"], 2)
    if "batched" in out:
        summary["batched_vs_plain"] = out["batched"]["batched_vs_plain"]
        summary["batched_msamples_s"] = out["batched"]["batched_msamples_s"]
    if "served" in out:
        k = out["served"]["knee"]
        summary["served_knee_vs_plain_loop"] = k["vs_plain_loop"]
        summary["served_knee_msamples_s"] = k["sustained_msamples_s"]
        summary["served_knee_latency_ms_p50_p99"] = k["latency_ms_p50_p99"]
        summary["served_max_vs_plain_loop"] = out["served"]["max_sustained"]["vs_plain_loop"]
    if "large" in out:
        summary["large_msamples_s"] = out["large"]["msamples_s"]
    out["summary"] = summary
    if args.detail_out and DETAIL:
        try:
            os.makedirs(os.path.dirname(os.path.abspath(args.detail_out)), exist_ok=True)
            with open(args.detail_out, "w") as f:
                json.dump(DETAIL, f, indent=1)
            out["detail_file"] = os.path.relpath(os.path.abspath(args.detail_out), ROOT)
        except OSError as e:
            log("detail tables not written: %s" % e)
        out["summary"] = out.pop("summary")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
