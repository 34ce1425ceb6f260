"""Parity at the benchmark's other shapes, synthetic device-resident state,
each checked chain by chain against the oracle (slices of the same state):

- two databases side by side on lean servers, ML-KEM-512 x 4M (1024 packed
  outputs, 32 chains, DB 8.6 GB) and ML-KEM-1024 x 250k (128 outputs, 4
  chains), one uneven 256-query batch split between them and answered in one
  pass on two streams;
- 32M ML-KEM-768 keys at dim0 65536 on a lean server (db_rows 65536, 352
  outputs, DB 94.5 GB: the last chain's columns lie 86 GB into it);
- the shape of the reference's published run, 5M ML-KEM-768 keys with
  8,388,608 buckets at dim0 65536 (64 outputs, 2 chains), both packing kernels,
  and a query from the library's KeywordClient at that shape.
"""
import sys

import pytest

import inspire_amd as IA
import oracle as O
from chain_oracle import chain_blobs, chain_of, split_response, stage_queries

pytestmark = pytest.mark.gpu
KEY = 32


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _free(torch, *servers):
    for s in servers:
        s.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_two_databases_uneven_batch_match_oracle():
    import torch
    shapes = [("ML-KEM-512 x 4M", 4_000_000, 800, 17, 201, [(0, 0), (100, 13), (200, 31)]),
              ("ML-KEM-1024 x 250k", 250_000, 1568, 18, 55, [(0, 3), (54, 1)])]
    dbs = []
    for name, recs, vs, seed, count, checks in shapes:
        plan = IA.plan_shape(recs, KEY, vs)
        say("%s: lean server, %d outputs, d %d, %d queries of the 256" % (
            name, plan["instances"], plan["interpolate_degree"], count))
        srv = IA.Server(plan["db_rows"], plan["instances"], plan["interpolate_degree"], lean=True)
        srv.fill_synthetic(seed)
        srv.set_max_batch(count)
        qs = [O.synth_query(seed * 1000 + i, plan["db_rows"], nq=2, small=(i % 7 == 3)) for i in range(count)]
        ser = [O.serialize_query(*q) for q in qs]
        dq, qlen, stride = stage_queries(torch, ser)
        blob = 2 * plan["c"] * O.RESP_BYTES
        dbs.append({"name": name, "plan": plan, "seed": seed, "srv": srv, "qs": qs, "ser": ser, "dq": dq,
                    "qlen": qlen, "stride": stride, "blob": blob, "count": count, "checks": checks,
                    "out": torch.zeros(count * blob, dtype=torch.uint8, device="cuda"),
                    "stream": torch.cuda.Stream()})
    assert sum(db["count"] for db in dbs) == 256
    torch.cuda.synchronize()
    say("one pass: both shares on their own streams")
    for db in dbs:
        db["srv"].respond_batch_device(db["dq"], db["stride"], db["qlen"], 2, db["count"], db["out"], db["blob"],
                                       db["stream"].cuda_stream)
    torch.cuda.synchronize()
    for db in dbs:
        out = db["out"].cpu().numpy().tobytes()
        blob, half, plan = db["blob"], db["blob"] // 2, db["plan"]
        say("%s: every query against its single-query answer" % db["name"])
        for i in range(db["count"]):
            s = split_response(db["srv"].respond_serialized(db["ser"][i]))
            assert out[i * blob:(i + 1) * blob] == s[0] + s[1], (db["name"], i)
        for i, w in db["checks"]:
            say("%s: query %d chain %d against the oracle" % (db["name"], i, w))
            exp = chain_blobs(db["seed"], plan["db_rows"], plan["interpolate_degree"], w, *db["qs"][i])
            mine = out[i * blob:(i + 1) * blob]
            for h in range(2):
                assert chain_of(mine[h * half:(h + 1) * half], w) == exp[h], (db["name"], i, w, h)
    for db in dbs:
        del db["dq"], db["out"]
    _free(torch, *[db["srv"] for db in dbs])


def test_32m_keys_dim0_65536_match_oracle():
    import torch
    plan = IA.plan_shape(32_000_000, KEY, 1184, dim0=65536)
    rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    assert (rows, inst, d) == (65536, 352, 32)
    say("32M keys: lean server, db_rows %d, %d outputs (DB %.1f GB)" % (rows, inst, plan["db_bytes"] / 1e9))
    srv = IA.Server(rows, inst, d, lean=True)
    srv.fill_synthetic(19)
    q = O.synth_query(1901, rows, nq=2)
    resp = split_response(srv.respond_serialized(O.serialize_query(*q)))
    c = inst // d
    assert len(resp) == 2 and all(len(b) == c * O.RESP_BYTES for b in resp)
    for w in (0, c - 1):
        say("32M keys: single query, chain %d against the oracle" % w)
        exp = chain_blobs(19, rows, d, w, *q)
        for h in range(2):
            assert chain_of(resp[h], w) == exp[h], (w, h)
    # a batch of 3 against the single-query answers, one chain against the oracle
    say("32M keys: batch of 3")
    qs = [O.synth_query(1910 + i, rows, nq=2, small=(i == 1)) for i in range(3)]
    ser = [O.serialize_query(*x) for x in qs]
    dq, qlen, stride = stage_queries(torch, ser)
    srv.set_max_batch(3)
    blob = 2 * c * O.RESP_BYTES
    dout = torch.zeros(3 * blob, dtype=torch.uint8, device="cuda")
    srv.respond_batch_device(dq, stride, qlen, 2, 3, dout, blob)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    for i in range(3):
        s = split_response(srv.respond_serialized(ser[i]))
        assert out[i * blob:(i + 1) * blob] == s[0] + s[1], i
    exp = chain_blobs(19, rows, d, 6, *qs[2])
    mine = out[2 * blob:3 * blob]
    for h in range(2):
        assert chain_of(mine[h * (blob // 2):(h + 1) * (blob // 2)], 6) == exp[h], h
    del dq, dout
    _free(torch, srv)


def test_published_5m_shape_match_oracle():
    # perf.md:36-51: 5,000,000 ML-KEM-768 keys, buckets 8,388,608, dim0 65536.
    import torch
    plan = IA.plan_shape(5_000_000, KEY, 1184, buckets=8_388_608, dim0=65536)
    rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    assert (rows, plan["db_cols"], inst, plan["c"]) == (65536, 131072, 64, 2)
    assert plan["peak_bytes"] == 32_205_963_264  # perf.md:55
    say("5M keys (published shape): db_rows %d, %d outputs" % (rows, inst))
    srv = IA.Server(rows, inst, d)
    srv.fill_synthetic(5)
    q = O.synth_query(501, rows, nq=2)
    ser = O.serialize_query(*q)
    exp = [chain_blobs(5, rows, d, w, *q) for w in range(plan["c"])]
    for kind, name in ((IA.PACKING_VALU, "valu"), (IA.PACKING_MFMA, "mfma")):
        say("5M keys: %s packing, both chains against the oracle" % name)
        srv.set_packing_kernel(kind)
        resp = split_response(srv.respond_serialized(ser))
        for w in range(plan["c"]):
            for h in range(2):
                assert chain_of(resp[h], w) == exp[w][h], (name, w, h)
    # the library's KeywordClient at this shape: a query row of 32 RLWE
    # encryptions per hash query, answered chain by chain as the oracle answers it
    say("5M keys: library client query (ternary secret) against the oracle")
    pp = dict(num_items=plan["num_items"], item_size_bits=plan["item_size_bits"], dim0=plan["dim0"],
              cuckoo_seed=O.CUCKOO_SEED, num_hashes=2, key_size=KEY, value_size=1184, entry_size=KEY + 1184,
              num_records=5_000_000)
    client = IA.KeywordClient(pp, hamming=O.HAMMING_WEIGHT, rng_seed=55)
    cq = client.query(bytes(range(KEY)))
    assert len(cq.data) == client.query_bytes() == 1_175_553   # 1 + (12,288 + 2 x (65,536 + 12,288)) x 7
    cy, cz, crows, cbodies = O.deserialize_query(cq.data, rows)
    resp = split_response(srv.respond_serialized(cq.data))
    for w in range(plan["c"]):
        exp = chain_blobs(5, rows, d, w, cy, cz, crows, cbodies)
        for h in range(2):
            assert chain_of(resp[h], w) == exp[h], ("client", w, h)
    _free(torch, srv)
