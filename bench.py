"""Benchmark: keyword-PIR respond throughput and latency on one MI355X.

Workload: 1,000,000 ML-KEM-768 keys (entry 1216 B) at the reference's default
shape (buckets 2,097,152, dim0 2048 -> db_rows 2048, db_cols 720,896,
352 packed outputs, interpolation degree 32, 11 fold chains).  Synthetic
device-resident state: encoded DB (2.95 GB) and InspiRING key-switch material
(35.4 GB condensed; 31.0 GB as the packed MFMA fragments the packing reads)
from the counter-based generator; the (32x smaller) cuckoo/encode
setup is not part of the timed path.  One step = one keyword query (2 hash
queries, the cuckoo num_hashes) answered end to end on the GPU from a
serialized query already resident in HBM: unpack, packing, INTT/CRT, GEMV,
RGSW fold, modulus switch and bit-pack.

Prints one JSON line (rank 0).  The cpu_baseline leg times the CPU oracle on a
bounded sample of the same workload and scales it to one keyword query.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inspire-rs_amd"))

METRIC = "PIR queries/sec + p50/p99 respond latency, 1M keys, batched vs single-query"
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA
HBM_PEAK_GBS = 8000.0
RECORDS, KEY, VALUE = 1_000_000, 32, 1184


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_queries(IA, torch, db_rows, count, seed, nq=2, return_host=False):
    """Serialized synthetic keyword queries (uniform residues, as from a client
    with a real ternary key), staged back to back in one device buffer."""
    import numpy as np
    N = IA.N
    blobs = []
    for i in range(count):
        s = seed * 1000 + i
        y = torch.empty(3 * N, dtype=torch.int64, device="cuda")
        z = torch.empty(3 * N, dtype=torch.int64, device="cuda")
        r = torch.empty(nq * db_rows, dtype=torch.int64, device="cuda")
        b = torch.empty(nq * 12 * N, dtype=torch.int64, device="cuda")
        IA.synth_fill(y, s, IA.STREAM_Y, 3 * N, IA.KIND_CONDENSED)
        IA.synth_fill(z, s, IA.STREAM_Z, 3 * N, IA.KIND_CONDENSED)
        IA.synth_fill(r, s, IA.STREAM_ROW, nq * db_rows, IA.KIND_CONDENSED)
        IA.synth_fill(b, s, IA.STREAM_BODY, nq * 12 * N, IA.KIND_CRT)
        torch.cuda.synchronize()
        hv = lambda t: t.cpu().numpy().view(np.uint64)
        blobs.append(IA.serialize_query(hv(y).reshape(3, N), hv(z).reshape(3, N),
                                        hv(r).reshape(nq, db_rows), hv(b).reshape(nq, 6, 2, N)))
    qlen = len(blobs[0])
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * count, dtype=np.uint8)
    for i, bl in enumerate(blobs):
        host[i * stride:i * stride + qlen] = np.frombuffer(bl, dtype=np.uint8)
    dev = torch.from_numpy(host).cuda()
    if return_host:
        return dev, qlen, stride, blobs
    return dev, qlen, stride


def serve_poisson(IA, eng, blobs, rate, n, seed):
    """Open-loop Poisson arrivals at `rate` q/s through the continuous-batching
    engine: a submitter thread sleeps to each arrival time and submits; a waiter
    collects each query's submit-to-ready latency (measured in the library)."""
    import queue
    import threading
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    tickets = queue.Queue()
    lat = []
    b0 = eng.stats()

    def waiter():
        while True:
            t = tickets.get()
            if t is None:
                return
            lat.append(eng.wait(t)[1])

    w = threading.Thread(target=waiter)
    w.start()
    dropped = 0
    t0 = time.perf_counter()
    for i in range(n):
        dt = t0 + arrivals[i] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        try:
            tickets.put(eng.submit(blobs[i % len(blobs)]))
        except IA.InspireError:
            dropped += 1
    tickets.put(None)
    w.join()
    elapsed = time.perf_counter() - t0
    b1 = eng.stats()
    lat.sort()
    pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
    nb = b1["batches"] - b0["batches"]
    return {"offered_qps": rate, "achieved_qps": len(lat) / elapsed, "queries": len(lat), "dropped": dropped,
            "p50_ms": pct(0.5), "p99_ms": pct(0.99), "mean_ms": sum(lat) / len(lat),
            "mean_batch": (b1["queries"] - b0["queries"]) / max(nb, 1)}


def bench_mixed(IA, torch, rounds=12, seed=37):
    """Two databases resident side by side on lean servers (fragment-ordered
    bold_t only): ML-KEM-512 x 4M (800-byte keys) and ML-KEM-1024 x 250k
    (1568-byte keys), synthetic state.  Each round draws a batch of 1-256
    keyword queries, splits it at random between the two databases and runs
    both shares in the same pass, each on its own stream.  Per database: the
    algorithmic bytes of its share (its DB once, its bold_t once) over the
    pass time."""
    import numpy as np
    rng = np.random.default_rng(seed)
    dbs = []
    for name, recs, vs in (("ML-KEM-512 x 4M", 4_000_000, 800), ("ML-KEM-1024 x 250k", 250_000, 1568)):
        plan = IA.plan_shape(recs, KEY, vs)
        t0 = time.time()
        srv = IA.Server(plan["db_rows"], plan["instances"], plan["interpolate_degree"], lean=True)
        srv.fill_synthetic(17)
        srv.set_max_batch(256)
        srv.set_graphs(False)
        dq, qlen, stride = make_queries(IA, torch, plan["db_rows"], 8, 29)
        dq = dq.repeat(32)[:256 * stride].contiguous()
        blob = 2 * plan["c"] * IA.RESP_BYTES
        out = torch.empty(256 * blob, dtype=torch.uint8, device="cuda")
        dbs.append({"name": name, "plan": plan, "srv": srv, "dq": dq, "qlen": qlen, "stride": stride, "blob": blob,
                    "out": out, "stream": torch.cuda.Stream(), "queries": 0, "bytes": 0.0})
        log("mixed: %s resident in %.1fs (instances %d)" % (name, time.time() - t0, plan["instances"]))

    def run(counts):
        for db, n in zip(dbs, counts):
            if n:
                db["srv"].respond_batch_device(db["dq"], db["stride"], db["qlen"], 2, n, db["out"], db["blob"],
                                               db["stream"].cuda_stream)
        torch.cuda.synchronize()

    run([64, 64])  # warm-up
    passes = []
    for r in range(rounds):
        total = int(rng.integers(1, 257))
        n0 = int(rng.binomial(total, 0.5))
        counts = [n0, total - n0]
        t0 = time.perf_counter()
        run(counts)
        dt = time.perf_counter() - t0
        passes.append({"queries": counts, "ms": dt * 1e3})
        for db, n in zip(dbs, counts):
            if n:
                db["queries"] += n
                # the DB once and bold_t once, 28 bits per residue as the lean
                # server's packed fragments hold it
                db["bytes"] += db["plan"]["db_bytes"] + db["plan"]["bold_t_bytes"] * 56 // 64
    wall = sum(x["ms"] for x in passes) / 1e3
    res = {"workload": "two lean servers on one GPU; batches of 1-256 keyword queries split at random between them, "
                       "both shares in one pass on two streams",
           "rounds": passes, "qps": sum(sum(x["queries"]) for x in passes) / wall, "dbs": []}
    for db in dbs:
        res["dbs"].append({"db": db["name"], "instances": db["plan"]["instances"], "queries": db["queries"],
                           "algorithmic_GBps": db["bytes"] / wall / 1e9,
                           "db_bytes": db["plan"]["db_bytes"],
                           "bold_t_bytes": db["plan"]["bold_t_bytes"] * 56 // 64})
        db["srv"].close()
    log("mixed: %.0f q/s over %d passes; %s" % (res["qps"], rounds, ", ".join(
        "%s %.0f GB/s" % (d["db"], d["algorithmic_GBps"]) for d in res["dbs"])))
    return res


def bench_32m(IA, torch, rates=(100, 300, 600, 700), serve_seconds=2.0):
    """32M ML-KEM-768 keys at dim0 65536 (DB 94.5 GB, bold_t 35.4 GB) on a lean
    server, synthetic state: single-query respond, batched respond and
    continuous-batching service under Poisson arrivals."""
    plan = IA.plan_shape(32_000_000, KEY, VALUE, dim0=65536)
    db_rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    t0 = time.time()
    srv = IA.Server(db_rows, inst, d, lean=True)
    srv.fill_synthetic(19)
    log("32m: state resident in %.1fs (db_rows %d, instances %d)" % (time.time() - t0, db_rows, inst))
    dq, qlen, stride, blobs = make_queries(IA, torch, db_rows, 4, 31, return_host=True)
    blob = 2 * plan["c"] * IA.RESP_BYTES
    out = torch.empty(64 * blob, dtype=torch.uint8, device="cuda")
    outs = [out[:blob], out[blob:2 * blob]]
    sts = [torch.cuda.Stream() for _ in range(2)]
    steps = 10

    def timed(in_flight):
        """ms per single-query respond with `in_flight` requests on their own streams."""
        for k in range(2):
            srv.respond_device(dq.data_ptr() + (k % 4) * stride, qlen, 2, outs[k % in_flight],
                               sts[k % in_flight].cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            srv.respond_device(dq.data_ptr() + (k % 4) * stride, qlen, 2, outs[k % in_flight],
                               sts[k % in_flight].cuda_stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    serial_ms = timed(1)
    single_ms = timed(2)
    res = {"workload": "32M ML-KEM-768 keys, dim0 65536, lean server",
           "db_rows": db_rows, "db_cols": plan["db_cols"], "instances": inst, "db_bytes": plan["db_bytes"],
           "bold_t_bytes": plan["bold_t_bytes"] * 56 // 64,  # packed fragments, 28 bits per residue
           "single_query": {"ms": single_ms, "qps": 1e3 / single_ms, "in_flight": 2,
                            "serial_ms": serial_ms, "serial_qps": 1e3 / serial_ms},
           "batched": {}, "serving": []}
    log("32m: single query %.2f ms (%.1f q/s) with two in flight, %.2f ms serial" %
        (single_ms, 1e3 / single_ms, serial_ms))
    srv.set_max_batch(64)
    for B in (8, 32, 64):
        bq = dq.repeat(B // 4 + 1)[:B * stride].contiguous()
        srv.respond_batch_device(bq, stride, qlen, 2, B, out, blob, sts[0].cuda_stream)
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            srv.respond_batch_device(bq, stride, qlen, 2, B, out, blob, sts[0].cuda_stream)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        res["batched"][str(B)] = {"ms_per_batch": ms, "qps": B * 1e3 / ms, "speedup_vs_single_query": B * single_ms / ms}
        log("32m: batch %d: %.1f ms, %.0f q/s" % (B, ms, B * 1e3 / ms))
        del bq
    eng = IA.Engine(srv, nq=2, max_batch=64, window_ms=0.0, capacity=256)
    for rate in rates:
        n = int(min(1200, max(50, rate * serve_seconds)))
        r = serve_poisson(IA, eng, blobs, float(rate), n, 41)
        res["serving"].append(r)
        log("32m: serve %.0f q/s offered: %.1f q/s done, p50 %.1f ms, p99 %.1f ms, mean batch %.1f, dropped %d" %
            (rate, r["achieved_qps"], r["p50_ms"], r["p99_ms"], r["mean_batch"], r["dropped"]))
    eng.close()
    srv.close()
    return res


def bench_published_5m(IA, torch, steps=20):
    """The shape of the reference's one published run (perf.md:36-64): 5M
    ML-KEM-768 keys, buckets 8,388,608, dim0 65536 (db_rows 65536, 131,072
    columns, 64 outputs, 2 chains), synthetic state on a lean server.  Single
    keyword-query respond (two in flight, as the headline), a batch of 64, and
    the device setup of that shape (DB encode + offline precomputation).  The
    published 1472.4 ms per query (perf.md:64) is context from other hardware,
    not a same-node comparison."""
    plan = IA.plan_shape(5_000_000, KEY, VALUE, buckets=8_388_608, dim0=65536)
    rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    t0 = time.time()
    srv = IA.Server(rows, inst, d, lean=True)
    srv.fill_synthetic(5)
    log("5m: state resident in %.1fs (db_rows %d, instances %d)" % (time.time() - t0, rows, inst))
    dq, qlen, stride = make_queries(IA, torch, rows, 4, 43)
    blob = 2 * plan["c"] * IA.RESP_BYTES
    outs = [torch.empty(blob, dtype=torch.uint8, device="cuda") for _ in range(2)]
    sts = [torch.cuda.Stream() for _ in range(2)]
    for k in range(4):
        srv.respond_device(dq.data_ptr() + (k % 4) * stride, qlen, 2, outs[k % 2], sts[k % 2].cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        srv.respond_device(dq.data_ptr() + (k % 4) * stride, qlen, 2, outs[k % 2], sts[k % 2].cuda_stream)
    torch.cuda.synchronize()
    single_ms = (time.perf_counter() - t0) * 1e3 / steps
    res = {"workload": "5M ML-KEM-768 keys, buckets 8,388,608, dim0 65536 (perf.md:36-51), lean server, synthetic",
           "db_rows": rows, "db_cols": plan["db_cols"], "instances": inst, "chains": plan["c"],
           "single_query": {"ms_per_query": single_ms, "qps": 1e3 / single_ms, "in_flight": 2},
           "published": {"ms_per_query": 1472.4, "source": "perf.md:64",
                         "hardware": "one sm_90 GPU with 78.66 GiB free (perf.md:56), other hardware"}}
    log("5m: single query %.3f ms (%.1f q/s); published 1472.4 ms" % (single_ms, 1e3 / single_ms))
    B = 64
    srv.set_max_batch(B)
    bq = dq.repeat(B // 4)[:B * stride].contiguous()
    bout = torch.empty(B * blob, dtype=torch.uint8, device="cuda")
    srv.respond_batch_device(bq, stride, qlen, 2, B, bout, blob, sts[0].cuda_stream)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        srv.respond_batch_device(bq, stride, qlen, 2, B, bout, blob, sts[0].cuda_stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    res["batched"] = {"B": B, "ms_per_batch": ms, "qps": B * 1e3 / ms}
    log("5m: batch %d: %.1f ms, %.0f q/s" % (B, ms, B * 1e3 / ms))
    del bq, bout
    srv.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    # device setup of this shape: encode of a raw DB, then the offline precomputation
    ncell = rows * plan["db_cols"]
    raw = torch.empty(ncell, dtype=torch.int16, device="cuda")
    IA.synth_fill(raw, 53, IA.STREAM_DB, ncell, IA.KIND_U16)
    outs_pre = [torch.empty(n, dtype=torch.int64, device="cuda")
                for n in (inst * 3069 * IA.N, inst * 3069 * IA.N, inst * 3 * IA.N, inst * IA.N)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    IA.encode_db_device(raw, raw, rows, plan["db_cols"], d)
    torch.cuda.synchronize()
    enc_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    IA.setup_offline_device(raw, rows, inst, *outs_pre)
    off_ms = (time.perf_counter() - t0) * 1e3
    res["setup"] = {"encode_ms": enc_ms, "offline_ms": off_ms,
                    "published_init_s": 264, "published_init_source": "perf.md:61 (dataset load, cuckoo table, "
                                                                       "encode and precomputation)"}
    log("5m: setup encode %.1f ms, offline precomputation %.1f ms (published init 264 s)" % (enc_ms, off_ms))
    del raw, outs_pre
    return res


def cpu_baseline(plan, out_samples=32, fold_chains=2):
    """Times the CPU oracle (restatement of the reference's respond path) on a
    bounded sample and scales it to one keyword query (2 hash queries)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    N, inst, d, rows = O.N, plan["instances"], plan["interpolate_degree"], plan["db_rows"]
    J = min(out_samples, inst)
    seed = 11
    bt = O.synth(seed, O.STREAM_BT, J * O.NINNER * N, O.KIND_CONDENSED)
    btb = O.synth(seed, O.STREAM_BTB, J * O.NINNER * N, O.KIND_CONDENSED)
    bth = O.synth(seed, O.STREAM_BTH, J * 3 * N, O.KIND_CONDENSED)
    db = O.synth(seed, O.STREAM_DB, J * N * rows, O.KIND_U16)
    y, z, qrows, bodies = O.synth_query(seed, rows, nq=2)
    t0 = time.perf_counter()
    S = O.packing(bt, btb, bth, y, z, J)
    t_pack = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.post(S)
    t_post = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h in range(2):
        O.gemv(qrows[h], db, rows, J * N)
    t_gemv = time.perf_counter() - t0
    neg_a, _ = O.rgsw_neg_a_ntt()
    pa = O.synth(seed, O.STREAM_AHAT, d * N, O.KIND_MODQ).reshape(d, N)
    pb = O.synth(seed + 1, O.STREAM_AHAT, d * N, O.KIND_MODQ).reshape(d, N)
    t0 = time.perf_counter()
    for _ in range(fold_chains):
        O.switch_and_keep(O.fold_chain(pa, pb, neg_a, bodies[0]))
    t_fold = (time.perf_counter() - t0) / fold_chains
    scale = inst / J
    per_query = (t_pack + t_post + t_gemv) * scale + t_fold * 2 * plan["c"]
    # DB encode: 64 rows of one column group, scaled to the whole DB.
    er = 64
    raw = O.synth(seed, O.STREAM_DB, er * d * N, O.KIND_U16)
    t0 = time.perf_counter()
    O.encode(raw, er, d * N, d)
    t_enc = (time.perf_counter() - t0) * (rows / er) * plan["c"]
    # Offline precomputation: one output polynomial (2048 columns), scaled.
    t0 = time.perf_counter()
    O.setup_precomp(O.setup_atilde(db, rows, 0, N), 1)
    t_setup = (time.perf_counter() - t0) * inst
    sample = ("packing+INTT for %d of %d outputs, GEMV of %d of %d columns x 2 hash queries, "
              "fold+modswitch of %d of %d chains; linearly scaled to one keyword query"
              % (J, inst, J * N, plan["db_cols"], fold_chains, 2 * plan["c"]))
    host = O.host_info()
    return {"value": 1.0 / per_query, "unit": "queries/s", "cores": O.num_threads(), "kind": "port",
            "sample": sample, "seconds_per_query": per_query,
            "host": {"cpu_model": host["cpu_model"], "hw_threads": host["hw_threads"],
                     "ram_gb": round(host["ram_bytes"] / 1e9, 1), "avx512": host["avx512"],
                     "oracle_build": "gcc -O3 -fopenmp, " + host["oracle_isa"],
                     "note": "the reference's CPU GEMV is hand-written AVX-512 (gpu/kernel.rs:139-237); the "
                             "oracle is plain C vectorised by the compiler for the ISA named here"},
            "stage_s": {"packing": t_pack * scale, "post": t_post * scale, "gemv": t_gemv * scale,
                        "fold": t_fold * 2 * plan["c"]},
            "encode_s": t_enc, "encode_sample": "64 of %d rows of 1 of %d column groups, scaled" % (rows, plan["c"]),
            "offline_setup_s": t_setup, "offline_setup_sample": "1 of %d outputs (hint, prep-pack, precompute, collapse), scaled" % inst}


def profiled_traffic(kernel_prefix):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary for it exists."""
    path = os.path.join(ROOT, "profiles", "respond_1m_single_query_trace_and_hbm.md")
    if not os.path.exists(path):
        return None
    in_hbm = False
    for line in open(path):
        in_hbm = in_hbm or line.startswith("## HBM counters")
        cells = [c.strip() for c in line.split("|")]
        if in_hbm and len(cells) > 5 and cells[1].startswith(kernel_prefix):
            try:
                return float(cells[4]) + float(cells[5])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--queries", type=int, default=8, help="distinct resident queries cycled through")
    ap.add_argument("--latency-samples", type=int, default=400,
                    help="requests in the run the p50/p99 latencies come from")
    ap.add_argument("--slots", type=int, default=2, help="requests in flight (request slots / streams)")
    ap.add_argument("--batches", default="1,8,32,64,128", help="batch sizes for the batched respond ('' to skip)")
    ap.add_argument("--rates", default="100,500,1000,2000,3000,4000",
                    help="Poisson arrival rates (q/s) for the serving run ('' to skip)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrivals per rate span about this long")
    ap.add_argument("--no-encode", action="store_true", help="skip the DB encode and offline setup measurements")
    ap.add_argument("--no-setup", action="store_true", help="skip the offline setup measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the two-database (4M + 250k) configuration")
    ap.add_argument("--no-32m", action="store_true", help="skip the 32M-key dim0-65536 configuration")
    ap.add_argument("--rates-32m", default="100,300,600,700",
                    help="Poisson arrival rates (q/s) of the 32M-key service runs ('' to skip)")
    ap.add_argument("--no-5m", action="store_true", help="skip the 5M-key shape of the reference's published run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default batched kernels")
    ap.add_argument("--packing", choices=("mfma", "valu"), default="mfma",
                    help="single-query packing kernel: int8 MFMA over the fragment-ordered bold_t, or the VALU "
                         "kernel over the reference layout")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import inspire_amd as IA

    plan = IA.plan_shape(RECORDS, KEY, VALUE)
    db_rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    log("plan", {k: plan[k] for k in ("num_items", "db_rows", "db_cols", "instances", "interpolate_degree", "c")})
    srv = IA.Server(db_rows, inst, d)
    t0 = time.time()
    srv.fill_synthetic(11)
    log("state resident in %.1fs" % (time.time() - t0))
    if args.packing == "mfma":
        t0 = time.time()
        srv.set_packing_kernel(IA.PACKING_MFMA)
        log("fragment-ordered bold_t built in %.1fs" % (time.time() - t0))
    dq, qlen, stride = make_queries(IA, torch, db_rows, args.queries, 13)
    # One output buffer and one stream per request slot: requests alternate
    # between the streams, so one request's fold (22 workgroups) overlaps the
    # next request's packing.  Streams use the default priority.
    nslots = max(1, args.slots)
    srv.set_slots(nslots)
    outs = [torch.empty(2 * plan["c"] * IA.RESP_BYTES, dtype=torch.uint8, device="cuda") for _ in range(nslots)]
    streams = [torch.cuda.Stream() for _ in range(nslots)]
    assert all(st.cuda_stream for st in streams), "need non-null stream handles"

    def step(k):
        i = k % nslots
        srv.respond_device(dq.data_ptr() + (k % args.queries) * stride, qlen, 2, outs[i], streams[i].cuda_stream)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    log("warmup done")

    def timed(steps, nstreams):
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        t_begin = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_begin.record(streams[0])
        for st in streams[1:nstreams]:
            st.wait_event(t_begin)
        for k in range(steps):
            st = streams[k % nstreams]
            starts[k].record(st)
            i = k % nstreams
            srv.respond_device(dq.data_ptr() + (k % args.queries) * stride, qlen, 2, outs[i], st.cuda_stream)
            ends[k].record(st)
        for k in range(max(0, steps - nstreams), steps):
            streams[0].wait_event(ends[k])
        t_end.record(streams[0])
        torch.cuda.synchronize()
        total = t_begin.elapsed_time(t_end)
        lat = sorted(a.elapsed_time(b) for a, b in zip(starts, ends))
        pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
        return total, pct(0.5), pct(0.99)

    # serial: one request at a time on one stream
    ser_n = max(100, args.steps // 4)
    ser_ms, ser_p50, ser_p99 = timed(ser_n, 1)
    ser_qps = ser_n / (ser_ms / 1e3)
    log("serial: %.3f ms/query, %.1f q/s" % (ser_ms / ser_n, ser_qps))
    # pipelined over the slots: the timed region of the JSON line (exactly K steps)
    total_ms, _, _ = timed(args.steps, nslots)
    qps = args.steps / (total_ms / 1e3)
    log("timed %d steps over %d slots: %.3f ms/step, %.1f q/s" % (args.steps, nslots, total_ms / args.steps, qps))
    # p50/p99 of the same pipelined respond from a run long enough for a p99
    # (a K of 20 would make p99 the maximum of 20 latencies)
    lat_n = max(args.latency_samples, args.steps)
    lat_ms, p50, p99 = timed(lat_n, nslots)
    log("latency run: %d requests, %.1f q/s, p50 %.3f ms, p99 %.3f ms" % (lat_n, lat_n / (lat_ms / 1e3), p50, p99))

    # Per-stage kernel times (hipEvents between the kernels on the respond stream).
    srv.profile(True)
    for k in range(args.profile_steps):
        step(k)
    torch.cuda.synchronize()
    stage_ms, cnt = srv.profile_read()
    srv.profile(False)
    names = ["unpack", "packing", "post", "gemv", "fold"]
    stage = {n: stage_ms[i] / max(cnt, 1) for i, n in enumerate(names)}
    # Algorithmic bytes of the packing kernel: bold_t and bold_t_bar once.  The
    # MFMA kernel reads them bit-packed, 28 bits per residue (56 of the 64 bits
    # of a condensed element); the VALU kernel reads the condensed u64 words and
    # also bold_t_hat (the MFMA path adds those terms in the post-process).
    if args.packing == "mfma":
        pack_bytes = plan["bold_t_bytes"] * 56 // 64
    else:
        pack_bytes = plan["bold_t_bytes"] + inst * 3 * IA.N * 8
    pack_kernel = "k_packing_mfma<1, 44>" if args.packing == "mfma" else "k_packing<8, true>"
    gemv_bytes = plan["db_bytes"]
    pack_gbs = pack_bytes / (stage["packing"] / 1e3) / 1e9
    gemv_gbs = gemv_bytes / (stage["gemv"] / 1e3) / 1e9
    log("stages (ms):", {k: round(v, 4) for k, v in stage.items()})

    result = {
        "metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": total_ms / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "single-query keyword respond (2 hash queries), 1M ML-KEM-768 keys",
                   "records": RECORDS, "entry_bytes": KEY + VALUE, "buckets": plan["num_items"],
                   "dim0": plan["dim0"], "db_rows": db_rows, "db_cols": plan["db_cols"], "instances": inst,
                   "interpolate_degree": d, "chains": plan["c"], "hash_queries": 2,
                   "packing_kernel": args.packing,
                   "parallelism": "1 GPU"},
        "p50_ms": p50, "p99_ms": p99, "latency_samples": lat_n,
        "latency_note": "p50/p99: submit-to-done of each request (two in flight), from a separate run of "
                        "latency_samples requests; value and ms_per_step come from the K timed steps",
        "serial": {"qps": ser_qps, "ms_per_query": ser_ms / ser_n, "p50_ms": ser_p50, "p99_ms": ser_p99,
                   "samples": ser_n},
        "slots": nslots,
        "stage_ms": stage,
        "roofline": {"bound": "hbm", "kernel": pack_kernel + " (InspiRING key-switch inner product)",
                     "achieved": pack_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": pack_gbs / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("inspire::" + pack_kernel), "algorithmic_bytes": pack_bytes,
                     "gemv": {"achieved": gemv_gbs, "frac": gemv_gbs / HBM_PEAK_GBS, "algorithmic_bytes": gemv_bytes}},
    }
    # Batched respond: B keyword queries per pass over the DB and bold_t.
    batched = {}
    if args.batches:
        sizes = [int(x) for x in args.batches.split(",") if x]
        t0 = time.time()
        srv.set_max_batch(max(sizes))
        log("batch buffers + fragment-ordered bold_t built in %.1fs" % (time.time() - t0))
        # Autotuner: kernel variants for 64-query batches, then the batch-size sweep.
        if not args.no_autotune:
            tb = min(64, max(sizes))
            t0 = time.time()
            tune = srv.autotune(tb, nq=2, reps=3)
            tune["count"] = tb
            result["autotune"] = tune
            log("autotune (%.1fs): packing %s ms -> %d, gemv %s ms -> %d, fold %s ms -> %d, suggested batch %d" % (
                time.time() - t0, [round(x, 2) for x in tune["packing_ms"]], tune["packing_variant"],
                [round(x, 2) for x in tune["gemv_ms"]], tune["gemv_variant"], [round(x, 2) for x in tune["fold_ms"]],
                tune["fold_variant"], tune["suggested_batch"]))
        single_ms = total_ms / args.steps
        for B in sizes:
            bq, bqlen, bstride = make_queries(IA, torch, db_rows, min(B, args.queries), 17)
            bout = torch.empty(B * 2 * plan["c"] * IA.RESP_BYTES, dtype=torch.uint8, device="cuda")
            blob = 2 * plan["c"] * IA.RESP_BYTES
            # cycle the distinct queries over the B positions of the batch
            if B > args.queries:
                reps = (B + args.queries - 1) // args.queries
                bq = bq.repeat(reps)[:B * bstride].contiguous()
            st = streams[0]

            def run_batch():
                srv.respond_batch_device(bq, bstride, bqlen, 2, B, bout, blob, st.cuda_stream)

            run_batch()
            torch.cuda.synchronize()
            reps = max(2, min(20, int(4000 / (B * single_ms + 1))))
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
            ev[0].record(st)
            for r in range(reps):
                run_batch()
                ev[r + 1].record(st)
            torch.cuda.synchronize()
            per = sorted(ev[r].elapsed_time(ev[r + 1]) for r in range(reps))
            tot = ev[0].elapsed_time(ev[reps])
            qps_b = B * reps / (tot / 1e3)
            srv.profile(True)
            run_batch()
            torch.cuda.synchronize()
            bst, _ = srv.profile_read()
            srv.profile(False)
            batched[str(B)] = {"qps": qps_b, "ms_per_batch": tot / reps, "p50_batch_ms": per[len(per) // 2],
                               "p99_batch_ms": per[-1], "speedup_vs_single_query": qps_b / qps,
                               "stage_ms": {n: bst[i] for i, n in enumerate(names)}}
            log("  stages:", {n: round(bst[i], 3) for i, n in enumerate(names)})
            log("batch %d: %.2f ms/batch, %.1f q/s (%.1fx single)" % (B, tot / reps, qps_b, qps_b / qps))
            del bq, bout
    result["batched"] = batched
    # int8 MFMA rooflines of the batched GEMMs at the largest batch <= 64
    # (stage times from hipEvents on the batch stream).
    rb = [int(b) for b in batched if int(b) <= 64]
    if rb:
        B = max(rb)
        st_b = batched[str(B)]["stage_ms"]
        pack_ops = B * 2 * inst * IA.N * 6141 * 16 * 2          # limb-pair MACs x 2
        gemv_ops = B * 16 * plan["db_cols"] * db_rows * 2 * 2    # 16 rows/query, 2 byte planes, MAC x 2
        mk = lambda ops, ms, kern: {"bound": "mfma", "kernel": kern, "achieved": ops / (ms / 1e3) / 1e12,
                                    "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                                    "frac": ops / (ms / 1e3) / 1e12 / INT8_PEAK_TOPS, "ops": ops, "ms": ms}
        tune = result.get("autotune") or {}
        pv, gv = tune.get("packing_variant", 0), tune.get("gemv_variant", 0)
        gemv_kernel = "k_gemv_lds (LDS-shared fragments)" if gv == 0 else "k_gemv_wide (query fragments from L2)"
        result["batched_roofline"] = {"B": B, "packing": mk(pack_ops, st_b["packing"], "k_packing_wide variant %d" % pv),
                                      "gemv": mk(gemv_ops, st_b["gemv"], gemv_kernel),
                                      "clock_note": "peak is the 2.4 GHz figure; under this load the packing runs at "
                                                    "1.6 GHz and the GEMV at 1.8-1.9 GHz (SQ_BUSY_CYCLES per shader "
                                                    "engine), where their matrix cores are busy 61-64% and 54-57% of "
                                                    "the kernel: profiles/batched_b64_clock_and_matrix_busy.md"}
        log("batched MFMA at B=%d: packing %.0f TOP/s (%.1f%%), gemv %.0f TOP/s (%.1f%%)" % (
            B, result["batched_roofline"]["packing"]["achieved"], 100 * result["batched_roofline"]["packing"]["frac"],
            result["batched_roofline"]["gemv"]["achieved"], 100 * result["batched_roofline"]["gemv"]["frac"]))

    # Setup: the DB encode (encode_kernel.cu) of a raw DB of the same shape,
    # device to device, in place; one synchronous call (work buffer included).
    if not args.no_encode:
        ncell = db_rows * plan["db_cols"]
        raw = torch.empty(ncell, dtype=torch.int16, device="cuda")
        IA.synth_fill(raw, 31, IA.STREAM_DB, ncell, IA.KIND_U16)
        IA.encode_db_device(raw, raw, db_rows, plan["db_cols"], d)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        IA.encode_db_device(raw, raw, db_rows, plan["db_cols"], d)
        enc_s = time.perf_counter() - t0
        result["setup"] = {"encode_ms": enc_s * 1e3, "encode_db_bytes": 2 * ncell,
                           "encode_gbs": 2 * ncell * 2 / enc_s / 1e9,
                           "note": "raw DB read once and encoded DB written once (2 x 2.95 GB) over the call's wall time"}
        log("encode %.1f ms (%.0f GB/s of read+write DB bytes)" % (enc_s * 1e3, 4 * ncell / enc_s / 1e9))
        # Offline precomputation of the encoded DB (hint, prep-pack, InspiRING
        # precompute and collapse): bold_t, bold_t_bar, bold_t_hat, a_hat.
        if not args.no_setup:
            outs_pre = [torch.empty(n, dtype=torch.int64, device="cuda")
                        for n in (inst * 3069 * IA.N, inst * 3069 * IA.N, inst * 3 * IA.N, inst * IA.N)]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            IA.setup_offline_device(raw, db_rows, inst, *outs_pre)
            off_s = time.perf_counter() - t0
            result["setup"].update({"offline_ms": off_s * 1e3, "offline_instances": inst,
                                    "offline_note": "one synchronous call: 352 outputs x (2048 column NTTs, "
                                                    "4096 slot NTTs, two 1023-step collapse chains), work memory included"})
            log("offline precomputation %.1f ms" % (off_s * 1e3))
            del outs_pre
        del raw
        torch.cuda.empty_cache()

    # Serving: continuous batching under open-loop Poisson arrivals.
    serving = []
    if args.rates:
        _, _, _, sblobs = make_queries(IA, torch, db_rows, args.queries, 19, return_host=True)
        eng = IA.Engine(srv, nq=2, max_batch=64, window_ms=0.0, capacity=2048)
        log("engine max_batch %d" % eng.stats()["max_batch"])
        for rate in [float(x) for x in args.rates.split(",") if x]:
            n = int(min(3000, max(100, rate * args.serve_seconds)))
            r = serve_poisson(IA, eng, sblobs, rate, n, 23)
            serving.append(r)
            log("serve %.0f q/s offered: %.1f q/s done, p50 %.2f ms, p99 %.2f ms, mean batch %.1f, dropped %d" %
                (rate, r["achieved_qps"], r["p50_ms"], r["p99_ms"], r["mean_batch"], r["dropped"]))
        eng.close()
    result["serving"] = {"engine": "continuous batching, max_batch 64, no arrival window, second batch in flight when a full one waits",
                         "runs": serving}

    srv.close()
    del outs, dq
    torch.cuda.empty_cache()
    # The other configurations, each on its own (lean) servers after the 1M one
    # is gone: two databases side by side, and the 32M-key serving shape.
    for key, flag, fn in (("mixed_dbs", args.no_mixed, bench_mixed), ("serving_32m", args.no_32m, bench_32m),
                          ("published_5m", args.no_5m, bench_published_5m)):
        if flag:
            continue
        try:
            if fn is bench_32m:
                result[key] = fn(IA, torch, rates=[float(x) for x in args.rates_32m.split(",") if x])
            else:
                result[key] = fn(IA, torch)
        except Exception as e:  # a configuration that does not fit is reported, not fatal
            result[key] = {"error": str(e)}
            log("%s failed: %s" % (key, e))
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(plan)
        log("cpu baseline %.3f q/s" % result["cpu_baseline"]["value"])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
