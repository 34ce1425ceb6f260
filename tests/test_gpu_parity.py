"""GPU parity: every stage of the answer path and the whole respond, run through
the C ABI of libinspire_hip.so on gfx950 and compared bit for bit with the CPU
oracle (or with the committed golden digests of it)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

pytestmark = pytest.mark.gpu
N = O.N


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    return _t


def dev_from(torch, arr):
    a = np.ascontiguousarray(arr)
    if a.dtype == np.uint64:
        t = torch.from_numpy(a.view(np.int64))
    elif a.dtype == np.uint32:
        t = torch.from_numpy(a.view(np.int32))
    elif a.dtype == np.uint16:
        t = torch.from_numpy(a.view(np.int16))
    else:
        t = torch.from_numpy(a)
    return t.cuda()


def host(torch, t, dtype):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def empty(torch, n, dtype):
    tdt = {np.uint64: torch.int64, np.uint32: torch.int32, np.uint16: torch.int16, np.uint8: torch.uint8}[dtype]
    return torch.zeros(n, dtype=tdt, device="cuda")


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_synth_fill_matches_oracle(torch, kind):
    say("synth_fill kind", kind)
    n, start = 5000, 123456
    dt = np.uint16 if kind == 2 else np.uint64
    d = empty(torch, n, dt)
    IA.synth_fill(d, 77, 3, n, kind, start=start)
    assert (host(torch, d, dt) == O.synth(77, 3, n, kind, start=start)).all()


@pytest.mark.parametrize("db_rows,nq", [(2048, 2), (4096, 1)])
def test_unpack_query_matches_wire_format(torch, db_rows, nq):
    say("unpack", db_rows, nq)
    y, z, rows, bodies = O.synth_query(31, db_rows, nq=nq)
    buf = O.serialize_query(y, z, rows, bodies)
    dq = dev_from(torch, np.frombuffer(buf + bytes(16), dtype=np.uint8))
    dy, dz = empty(torch, 3 * N, np.uint64), empty(torch, 3 * N, np.uint64)
    dr, db = empty(torch, nq * db_rows, np.uint64), empty(torch, nq * 6 * 2 * N, np.uint32)
    IA.unpack_query_device(dq, nq, db_rows, dy, dz, dr, db)
    assert (host(torch, dy, np.uint64).reshape(3, N) == y).all()
    assert (host(torch, dz, np.uint64).reshape(3, N) == z).all()
    assert (host(torch, dr, np.uint64).reshape(nq, db_rows) == rows).all()
    assert (host(torch, db, np.uint32).reshape(nq, 6, 2, N) == bodies).all()


@pytest.mark.parametrize("db_rows,db_cols,nq,small", [(2048, 3000, 2, False), (4096, 257, 1, True),
                                                       (2048, 64, 2, False)])
def test_gemv_matches_oracle(torch, db_rows, db_cols, nq, small):
    say("gemv", db_rows, db_cols, nq, small)
    db = O.synth(41, O.STREAM_DB, db_rows * db_cols, O.KIND_U16)
    rows = O.synth(42, O.STREAM_ROW, nq * db_rows, O.KIND_SMALL if small else O.KIND_CONDENSED)
    dc = empty(torch, nq * db_cols, np.uint64)
    IA.gemv_device(dc, dev_from(torch, rows), nq, dev_from(torch, db), db_rows, db_cols)
    got = host(torch, dc, np.uint64).reshape(nq, db_cols)
    for h in range(nq):
        assert (got[h] == O.gemv(rows[h * db_rows:(h + 1) * db_rows], db, db_rows, db_cols)).all()


def test_gemv_extreme_values(torch):
    say("gemv extreme")
    db_rows, db_cols = 2048, 40
    db = np.full(db_rows * db_cols, 65534, dtype=np.uint16)
    db[::7] = 0
    a = np.full(db_rows, ((1 << 28) - 1) | (((1 << 28) - 1) << 32), dtype=np.uint64)
    dc = empty(torch, db_cols, np.uint64)
    IA.gemv_device(dc, dev_from(torch, a), 1, dev_from(torch, db), db_rows, db_cols)
    assert (host(torch, dc, np.uint64) == O.gemv(a, db, db_rows, db_cols)).all()


@pytest.mark.parametrize("db_rows,db_cols,nq,kind", [(2048, 3000, 2, "uniform"), (2048, 128, 1, "small"),
                                                       (4096, 200, 2, "uniform"), (65536, 40, 2, "extreme"),
                                                       (2048, 16, 2, "extreme")])
def test_gemv_mfma_matches_oracle(torch, db_rows, db_cols, nq, kind):
    say("gemv mfma", db_rows, db_cols, nq, kind)
    if kind == "extreme":
        db = np.full(db_rows * db_cols, 65534, dtype=np.uint16)
        db[::3] = 0
        db[1::5] = 255
        db[2::7] = 256
        rows = np.full(nq * db_rows, ((1 << 28) - 1) | (((1 << 28) - 1) << 32), dtype=np.uint64)
        rows[::11] = 0
    else:
        db = O.synth(43, O.STREAM_DB, db_rows * db_cols, O.KIND_U16)
        rows = O.synth(44, O.STREAM_ROW, nq * db_rows, O.KIND_SMALL if kind == "small" else O.KIND_CONDENSED)
    scratch = empty(torch, IA.gemv_mfma_scratch_bytes(db_rows), np.uint8)
    dc = empty(torch, nq * db_cols, np.uint64)
    IA.gemv_mfma_device(dc, dev_from(torch, rows), nq, dev_from(torch, db), db_rows, db_cols, scratch)
    got = host(torch, dc, np.uint64).reshape(nq, db_cols)
    for h in range(nq):
        assert (got[h] == O.gemv(rows[h * db_rows:(h + 1) * db_rows], db, db_rows, db_cols)).all(), h


@pytest.mark.parametrize("instances", [1, 2, 3, 8, 12])
def test_packing_matches_oracle(torch, instances):
    say("packing", instances)
    bt = O.synth(51, O.STREAM_BT, instances * O.NINNER * N, O.KIND_CONDENSED)
    btb = O.synth(51, O.STREAM_BTB, instances * O.NINNER * N, O.KIND_CONDENSED)
    bth = O.synth(51, O.STREAM_BTH, instances * 3 * N, O.KIND_CONDENSED)
    y, z, _, _ = O.synth_query(52, 2048, nq=1)
    if instances == 3:  # non-canonical 28-bit key residues from the wire
        y = y | np.uint64((1 << 28) - 1)
    dS = empty(torch, instances * 2 * N, np.uint32)
    IA.packing_device(dS, dev_from(torch, bt), dev_from(torch, btb), dev_from(torch, bth), dev_from(torch, y),
                      dev_from(torch, z), instances)
    got = host(torch, dS, np.uint32).reshape(instances, 2, N)
    exp = O.packing(bt, btb, bth, y, z, instances)
    assert (got == exp).all()


@pytest.mark.parametrize("count", [1, 5])
def test_post_matches_oracle(torch, count):
    say("post", count)
    S = O.synth(61, 7, count * 2 * N, O.KIND_CRT).reshape(count, 2, N)
    dP = empty(torch, count * N, np.uint64)
    IA.post_device(dP, dev_from(torch, S.astype(np.uint32)), count)
    assert (host(torch, dP, np.uint64).reshape(count, N) == O.post(S)).all()


@pytest.mark.parametrize("instances,d,nq", [(2, 2, 1), (4, 2, 2), (8, 8, 2), (6, 3, 1)])
def test_fold_matches_oracle(torch, instances, d, nq):
    say("fold", instances, d, nq)
    c = instances // d
    a_hat = O.synth(71, O.STREAM_AHAT, instances * N, O.KIND_MODQ)
    cvec = O.synth(72, 8, nq * instances * N, O.KIND_MODQ)
    P = O.synth(73, 9, instances * N, O.KIND_MODQ)
    bodies = O.synth(74, O.STREAM_BODY, nq * 6 * 2 * N, O.KIND_CRT)
    dout = empty(torch, nq * c * O.RESP_BYTES, np.uint8)
    IA.fold_device(dout, dev_from(torch, a_hat), dev_from(torch, cvec), dev_from(torch, P),
                   dev_from(torch, bodies.astype(np.uint32)), instances, d, nq)
    got = host(torch, dout, np.uint8).reshape(nq, c, O.RESP_BYTES)
    neg_a, _ = O.rgsw_neg_a_ntt()
    ah = a_hat.reshape(instances, N)
    Pm = P.reshape(instances, N)
    for h in range(nq):
        fb = cvec.reshape(nq, instances, N)[h] + Pm
        body = bodies.reshape(nq, 6, 2, N)[h]
        for w in range(c):
            ct = O.fold_chain(ah[w * d:(w + 1) * d], fb[w * d:(w + 1) * d], neg_a, body)
            assert got[h, w].tobytes() == O.switch_and_keep(ct), (h, w)


def test_fold_many_chains_matches_oracle(torch):
    # 512 chains in one launch select the 512-thread fold (primes in turn) of
    # the stateless entry point; a sample of chains against the oracle.
    instances, d, nq = 1024, 2, 1
    say("fold, 512 chains", instances, d, nq)
    c = instances // d
    a_hat = O.synth(81, O.STREAM_AHAT, instances * N, O.KIND_MODQ)
    cvec = O.synth(82, 8, nq * instances * N, O.KIND_MODQ)
    P = O.synth(83, 9, instances * N, O.KIND_MODQ)
    bodies = O.synth(84, O.STREAM_BODY, nq * 6 * 2 * N, O.KIND_CRT)
    dout = empty(torch, nq * c * O.RESP_BYTES, np.uint8)
    IA.fold_device(dout, dev_from(torch, a_hat), dev_from(torch, cvec), dev_from(torch, P),
                   dev_from(torch, bodies.astype(np.uint32)), instances, d, nq)
    got = host(torch, dout, np.uint8).reshape(nq, c, O.RESP_BYTES)
    neg_a, _ = O.rgsw_neg_a_ntt()
    ah = a_hat.reshape(instances, N)
    fb = cvec.reshape(nq, instances, N)[0] + P.reshape(instances, N)
    body = bodies.reshape(nq, 6, 2, N)[0]
    for w in (0, 1, 255, 256, 510, 511):
        ct = O.fold_chain(ah[w * d:(w + 1) * d], fb[w * d:(w + 1) * d], neg_a, body)
        assert got[0, w].tobytes() == O.switch_and_keep(ct), w


def _golden():
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", "respond_small.json")))["cases"]


@pytest.mark.parametrize("case", _golden(), ids=lambda c: "r%d_i%d_d%d" % (c["db_rows"], c["instances"], c["d"]))
def test_respond_serialized_matches_golden(torch, case):
    say("respond golden", case["db_rows"], case["instances"], case["d"], case["nq"])
    db_rows, inst, d, nq = case["db_rows"], case["instances"], case["d"], case["nq"]
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(case["state_seed"])
    stash = [bytes([i] * 1216) for i in range(3)]
    srv.set_stash(stash)
    y, z, rows, bodies = O.synth_query(case["query_seed"], db_rows, nq=nq, small=case["small"])
    q = O.serialize_query(y, z, rows, bodies)
    resp = srv.respond_serialized(q)
    blob = (inst // d) * O.RESP_BYTES
    assert len(resp) == 4 + nq * (4 + blob) + 4 + 3 * (4 + 1216)
    parts = [resp[8 + h * (4 + blob): 8 + h * (4 + blob) + blob] for h in range(nq)]
    assert [hashlib.sha256(p).hexdigest() for p in parts] == case["sha256"]
    # and the full serialization equals the oracle's wire format
    assert resp == O.serialize_response(parts, stash)
    srv.close()


def test_respond_serialized_matches_live_oracle(torch):
    say("respond live oracle 2048x4 d2")
    db_rows, inst, d = 2048, 4, 2
    st = O.synth_state(9, db_rows, inst)
    srv = IA.Server(db_rows, inst, d)
    srv.set_db(st[0])
    srv.set_precomp(st[1], st[2], st[3], st[4])
    y, z, rows, bodies = O.synth_query(10, db_rows, nq=2)
    resp = srv.respond_serialized(O.serialize_query(y, z, rows, bodies))
    exp = O.respond_raw(st[0], db_rows, inst, d, st[1], st[2], st[3], st[4], y, z, rows, bodies)
    assert resp == O.serialize_response(exp, [])
    # repeat: the resident state is unchanged by a respond
    assert srv.respond_serialized(O.serialize_query(y, z, rows, bodies)) == resp
    srv.close()


def test_respond_gemv_kernels_agree(torch):
    say("respond valu vs mfma gemv")
    db_rows, inst, d = 4096, 4, 4
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(12)
    y, z, rows, bodies = O.synth_query(13, db_rows, nq=2)
    q = O.serialize_query(y, z, rows, bodies)
    srv.set_gemv_kernel(IA.GEMV_MFMA)
    a = srv.respond_serialized(q)
    srv.set_gemv_kernel(IA.GEMV_VALU)
    b = srv.respond_serialized(q)
    assert a == b
    srv.close()


def test_respond_edge_cases(torch):
    say("respond edge cases")
    db_rows, inst, d = 2048, 2, 2
    srv = IA.Server(db_rows, inst, d)
    y, z, rows, bodies = O.synth_query(11, db_rows, nq=2)
    q = O.serialize_query(y, z, rows, bodies)
    with pytest.raises(IA.InspireError) as e:   # no state loaded yet
        srv.respond_serialized(q)
    assert e.value.code == IA.ERR_STATE
    srv.fill_synthetic(5)
    with pytest.raises(IA.InspireError) as e:   # truncated
        srv.respond_serialized(q[:-3])
    assert e.value.code == IA.ERR_INVALID
    with pytest.raises(IA.InspireError):        # empty
        srv.respond_serialized(b"")
    # zero hash queries: only the counts and the stash come back
    q0 = O.serialize_query(y, z, rows[:0], bodies[:0])
    assert srv.respond_serialized(q0) == O.serialize_response([], [])
    # trailing bytes after a well-formed query are ignored (wire.rs:618-705)
    assert srv.respond_serialized(q + b"\x00" * 9) == srv.respond_serialized(q)
    srv.close()


def test_legacy_gemv_query(torch):
    say("legacy gemv")
    import ctypes
    L = IA.lib()
    db_rows, db_cols = 2048, 1000
    db = O.synth(81, O.STREAM_DB, db_rows * db_cols, O.KIND_U16)
    a = O.synth(82, O.STREAM_ROW, db_rows, O.KIND_CONDENSED)
    d_db = ctypes.c_void_p()
    L.gpu_gemv_upload_db(db.ctypes.data_as(ctypes.c_void_p), ctypes.byref(d_db), ctypes.c_size_t(db.nbytes))
    assert d_db.value
    c = np.zeros(db_cols, dtype=np.uint64)
    k = O.constants()
    u = ctypes.c_uint64
    L.gpu_gemv_query(c.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p), d_db, db_rows, db_cols,
                     u(O.Q0), u(O.Q1), u(k["cr1_q0"]), u(k["cr1_q1"]), u(k["mod0_inv_mod1"]),
                     u(k["mod1_inv_mod0"]), u(O.Q), u(k["cr0_q"]), u(k["cr1_q"]))
    L.gpu_gemv_free_db(d_db)
    assert (c == O.gemv(a, db, db_rows, db_cols)).all()


def test_legacy_rotation_and_packing(torch):
    say("legacy rotation + packing")
    import ctypes
    L = IA.lib()
    vp = ctypes.c_void_p
    inst = 2
    tables = np.concatenate([O.automorph_table(t) for t in range(1, 2 * N, 2)]).astype(np.uint32)
    gen_pows = np.array([pow(5, i, 2 * N) for i in range(O.NROT)], dtype=np.uint32)
    L.gpu_rotation_upload_tables(tables.ctypes.data_as(vp), gen_pows.ctypes.data_as(vp), N, O.NROT, N)
    L.gpu_packing_online_ensure_key_buffers(O.NINNER, 3, N)
    y, z, _, _ = O.synth_query(83, 2048, nq=1)
    L.gpu_rotation_expand_keys(vp(L.gpu_packing_online_get_d_y_all()), vp(L.gpu_packing_online_get_d_y_bar_all()),
                               y.ctypes.data_as(vp), O.NROT, 3, N)
    L.gpu_packing_online_upload_z_body(z.ctypes.data_as(vp), 3, N)
    bt = O.synth(84, O.STREAM_BT, inst * O.NINNER * N, O.KIND_CONDENSED)
    btb = O.synth(84, O.STREAM_BTB, inst * O.NINNER * N, O.KIND_CONDENSED)
    bth = O.synth(84, O.STREAM_BTH, inst * 3 * N, O.KIND_CONDENSED)
    L.gpu_packing_online_upload_precomp(bt.ctypes.data_as(vp), btb.ctypes.data_as(vp), bth.ctypes.data_as(vp),
                                        inst, O.NINNER, 3, N)
    out = np.zeros(inst * 2 * N, dtype=np.uint64)
    k = O.constants()
    u = ctypes.c_uint64
    L.gpu_packing_online_compute(out.ctypes.data_as(vp), inst, O.NROT, 3, N, 128, u(O.Q0), u(O.Q1),
                                 u(k["cr1_q0"]), u(k["cr1_q1"]))
    L.gpu_packing_online_free()
    L.gpu_rotation_free()
    assert (out.reshape(inst, 2, N) == O.packing(bt, btb, bth, y, z, inst)).all()


# ---------------------------------------------------------------------------
# MFMA packing and batched respond
# ---------------------------------------------------------------------------

def _queries(db_rows, seeds, nq=2, small=False):
    out = []
    for s in seeds:
        y, z, rows, bodies = O.synth_query(s, db_rows, nq=nq, small=small)
        out.append((O.serialize_query(y, z, rows, bodies), (y, z, rows, bodies)))
    return out


def _split_response(resp):
    """The hash queries' blobs of a serialized keyword response (wire.rs:708-726)."""
    n = int.from_bytes(resp[:4], "little")
    out, o = [], 4
    for _ in range(n):
        ln = int.from_bytes(resp[o:o + 4], "little")
        out.append(resp[o + 4:o + 4 + ln])
        o += 4 + ln
    return out


@pytest.mark.parametrize("kind", [IA.PACKING_VALU, IA.PACKING_MFMA])
def test_single_query_requests_in_flight_equal_solo(torch, kind):
    # Nine single-query responds issued back to back on three streams over
    # three request slots, with no host synchronization between them (the
    # server orders the requests' GEMV and packing stages with an event and
    # reuses each slot after its previous request): every response equals the
    # same query answered alone.
    say("requests in flight, packing kind", kind)
    db_rows, inst, d, nq = 2048, 8, 4, 2
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(61)
    srv.set_packing_kernel(kind)
    srv.set_slots(3)
    qs = _queries(db_rows, [700 + i for i in range(9)])
    solo = [b"".join(_split_response(srv.respond_serialized(q))) for q, _ in qs]
    blob = nq * (inst // d) * O.RESP_BYTES
    assert all(len(x) == blob for x in solo)
    streams = [torch.cuda.Stream() for _ in range(3)]
    dqs = [dev_from(torch, np.frombuffer(q + bytes(16), dtype=np.uint8)) for q, _ in qs]
    outs = [empty(torch, blob, np.uint8) for _ in qs]
    torch.cuda.synchronize()
    for i, (q, _) in enumerate(qs):
        srv.respond_device(dqs[i], len(q), nq, outs[i], streams[i % 3].cuda_stream)
    torch.cuda.synchronize()
    for i in range(len(qs)):
        assert outs[i].cpu().numpy().tobytes() == solo[i], i
    srv.close()


@pytest.mark.parametrize("db_rows,inst,d", [(2048, 8, 8), (4096, 12, 4), (2048, 2, 1)])
def test_packing_mfma_respond_matches_valu(torch, db_rows, inst, d):
    say("packing mfma vs valu respond", db_rows, inst, d)
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(21)
    for q, _ in _queries(db_rows, [22, 23]):
        srv.set_packing_kernel(IA.PACKING_VALU)
        a = srv.respond_serialized(q)
        srv.set_packing_kernel(IA.PACKING_MFMA)
        b = srv.respond_serialized(q)
        assert a == b
    srv.close()


def test_packing_mfma_matches_live_oracle(torch):
    say("packing mfma vs oracle")
    db_rows, inst, d = 2048, 4, 2
    st = O.synth_state(31, db_rows, inst)
    srv = IA.Server(db_rows, inst, d)
    srv.set_db(st[0])
    srv.set_precomp(st[1], st[2], st[3], st[4])
    srv.set_packing_kernel(IA.PACKING_MFMA)
    q, (y, z, rows, bodies) = _queries(db_rows, [32], small=True)[0]
    exp = O.respond_raw(st[0], db_rows, inst, d, st[1], st[2], st[3], st[4], y, z, rows, bodies)
    assert srv.respond_serialized(q) == O.serialize_response(exp, [])
    srv.close()


# Client counts cover every packing configuration: 4- and 16-client tiles,
# 32- and 64-client wide blocks, and two 64-client groups (70); 96 instances
# give several N-tile groups with a partial last one.
@pytest.mark.parametrize("count,inst,d", [(1, 8, 4), (3, 8, 4), (5, 8, 4), (17, 8, 4), (20, 8, 4), (40, 8, 4),
                                          (70, 8, 4), (24, 96, 8), (40, 96, 8)])
def test_batch_results_equal_solo_results(torch, count, inst, d):
    # every query's response is the same alone and in the mix
    say("batch vs solo", count, inst, d)
    db_rows = 2048
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(41)
    srv.set_max_batch(max(32, count))
    qs = _queries(db_rows, [100 + i for i in range(count)], small=False)
    qlen = len(qs[0][0])
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * count, dtype=np.uint8)
    for i, (q, _) in enumerate(qs):
        host[i * stride:i * stride + qlen] = np.frombuffer(q, dtype=np.uint8)
    dq = dev_from(torch, host)
    blob = 2 * (inst // d) * O.RESP_BYTES
    dout = empty(torch, count * blob, np.uint8)
    srv.respond_batch_device(dq, stride, qlen, 2, count, dout, blob)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    for i, (q, _) in enumerate(qs):
        solo = srv.respond_serialized(q)
        parts = solo[8:8 + blob // 2] + solo[8 + blob // 2 + 4:8 + blob // 2 + 4 + blob // 2]
        assert out[i * blob:(i + 1) * blob] == parts, i
    # one query also against the oracle
    st = O.synth_state(41, db_rows, inst)
    y, z, rows, bodies = qs[0][1]
    exp = O.respond_raw(st[0], db_rows, inst, d, st[1], st[2], st[3], st[4], y, z, rows, bodies)
    assert out[:blob] == b"".join(exp)
    srv.close()


@pytest.mark.parametrize("count,inst,d", [(70, 8, 4), (64, 96, 8), (10, 4, 2)])
def test_batch_kernel_variants_agree_and_autotune(torch, count, inst, d):
    # Every packing / GEMV / fold variant gives the same bytes, and so does the
    # autotuner's choice.
    say("batch variants", count, inst, d)
    db_rows = 2048
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(43)
    srv.set_max_batch(count)
    qs = _queries(db_rows, [300 + i for i in range(count)], small=False)
    qlen = len(qs[0][0])
    stride = (qlen + 16 + 255) // 256 * 256
    hbuf = np.zeros(stride * count, dtype=np.uint8)
    for i, (q, _) in enumerate(qs):
        hbuf[i * stride:i * stride + qlen] = np.frombuffer(q, dtype=np.uint8)
    dq = dev_from(torch, hbuf)
    blob = 2 * (inst // d) * O.RESP_BYTES

    def run():
        dout = empty(torch, count * blob, np.uint8)
        srv.respond_batch_device(dq, stride, qlen, 2, count, dout, blob)
        torch.cuda.synchronize()
        return dout.cpu().numpy().tobytes()

    ref = run()
    # the library refuses variant numbers past its last one
    npack = 0
    while True:
        try:
            srv.set_batch_kernels(npack, 0, 0)
        except IA.InspireError:
            break
        npack += 1
    assert npack >= 5
    ngemv = 0
    while True:
        try:
            srv.set_batch_kernels(0, ngemv, 0)
        except IA.InspireError:
            break
        ngemv += 1
    assert ngemv >= 2
    for pv in range(npack):
        for gv in range(ngemv):
            for fv in range(3):
                srv.set_batch_kernels(pv, gv, fv)
                assert run() == ref, (pv, gv, fv)
    t = srv.autotune(count, nq=2, reps=2)
    say("autotune", t)
    # packing variants differ (and are timed) only for groups of 64 clients
    assert len(t["packing_ms"]) == (npack if count > 32 else 1)
    assert min(t["packing_ms"]) > 0 and min(t["fold_ms"]) > 0
    assert len(t["gemv_ms"]) == ngemv and min(t["gemv_ms"]) > 0
    assert t["suggested_batch"] in t["batch_qps"]
    assert run() == ref
    with pytest.raises(IA.InspireError):
        srv.set_batch_kernels(npack, 0, 0)
    srv.close()


def test_batch_graph_replays_equal_direct_launches(torch):
    # A batch key is launched directly on first use, captured on the second and
    # replayed after; each replay (with other work and synchronizations in
    # between, and the output cleared first) must equal the direct result.
    say("batch graphs")
    db_rows, inst, d, count = 2048, 96, 8, 24
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(47)
    srv.set_max_batch(40)
    qs = _queries(db_rows, [400 + i for i in range(count + 5)], small=False)
    qlen = len(qs[0][0])
    stride = (qlen + 16 + 255) // 256 * 256
    hbuf = np.zeros(stride * (count + 5), dtype=np.uint8)
    for i, (q, _) in enumerate(qs):
        hbuf[i * stride:i * stride + qlen] = np.frombuffer(q, dtype=np.uint8)
    dq = dev_from(torch, hbuf)
    blob = 2 * (inst // d) * O.RESP_BYTES
    dout = empty(torch, (count + 5) * blob, np.uint8)

    def run(n, q_off=0):
        dout.zero_()
        srv.respond_batch_device(dq[q_off * stride:], stride, qlen, 2, n, dout, blob)
        torch.cuda.synchronize()
        return dout[:n * blob].cpu().numpy().tobytes()

    srv.set_graphs(False)
    ref = run(count)
    ref5 = run(5, count)
    srv.set_graphs(True)
    for rep in range(5):
        assert run(count) == ref, rep
        assert srv.respond_serialized(qs[rep][0])  # other work on the server between replays
        assert run(5, count) == ref5, rep
    srv.set_batch_kernels(1, 1, 2)  # a new key: direct, captured, replayed
    for rep in range(3):
        assert run(count) == ref, ("variants", rep)
    srv.close()


# ---------------------------------------------------------------------------
# Continuous-batching engine
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("window_ms,max_batch", [(5.0, 8), (0.0, 32)])
def test_engine_responses_equal_solo_responses(torch, window_ms, max_batch):
    say("engine vs solo", window_ms, max_batch)
    import threading
    db_rows, inst, d = 2048, 8, 4
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(51)
    stash = [bytes([i]) * 16 for i in range(3)]
    srv.set_stash(stash)
    qparts = _queries(db_rows, [200 + i for i in range(24)])
    qs = [q for q, _ in qparts]
    solo = [srv.respond_serialized(q) for q in qs]
    eng = IA.Engine(srv, nq=2, max_batch=max_batch, window_ms=window_ms, capacity=64)
    tickets = [None] * len(qs)

    def submit(lo, hi):
        for i in range(lo, hi):
            tickets[i] = eng.submit(qs[i])

    th = [threading.Thread(target=submit, args=(k * 12, k * 12 + 12)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    got = []
    for i, t in enumerate(tickets):
        resp, lat = eng.wait(t)
        assert resp == solo[i], i
        assert lat > 0
        got.append(resp)
    st = eng.stats()
    assert st["queries"] == len(qs) and 1 <= st["batches"] <= len(qs)
    assert st["max_batch"] == max_batch
    eng.close()
    srv.close()
    # the engine's answers are the oracle's (responses and stash framing)
    state = O.synth_state(51, db_rows, inst)
    for i in (0, 13, 23):
        exp = O.respond_raw(state[0], db_rows, inst, d, *state[1:], *qparts[i][1])
        assert got[i] == O.serialize_response(exp, stash), i


def test_engine_rejects_bad_queries_and_full_queue(torch):
    say("engine errors")
    db_rows, inst, d = 2048, 8, 4
    srv = IA.Server(db_rows, inst, d)
    srv.fill_synthetic(52)
    q = _queries(db_rows, [300])[0][0]
    eng = IA.Engine(srv, nq=2, max_batch=4, window_ms=50.0, capacity=2)
    with pytest.raises(IA.InspireError):
        eng.submit(q[:100])                      # truncated
    with pytest.raises(IA.InspireError):
        eng.submit(bytes([1]) + q[1:])           # hash-query count differs
    t1, t2 = eng.submit(q), eng.submit(q)
    with pytest.raises(IA.InspireError):
        eng.submit(q)                            # capacity 2
    assert eng.wait(t1)[0] == eng.wait(t2)[0]
    with pytest.raises(IA.InspireError):
        eng.wait(t1)                             # already waited
    eng.close()
    srv.close()


def test_engines_on_two_databases(torch):
    # Two resident databases of different shapes, each behind its own engine on
    # the one GPU; queries to both are submitted interleaved from two threads.
    say("engines on two databases")
    import threading
    shapes = [(2048, 8, 4, 61), (4096, 12, 4, 62)]
    srvs, engs, qs, solo = [], [], [], []
    for rows, inst, d, seed in shapes:
        srv = IA.Server(rows, inst, d)
        srv.fill_synthetic(seed)
        q = [x for x, _ in _queries(rows, [seed * 10 + i for i in range(10)])]
        solo.append([srv.respond_serialized(x) for x in q])
        srvs.append(srv)
        qs.append(q)
    for srv in srvs:
        engs.append(IA.Engine(srv, nq=2, max_batch=8, window_ms=2.0, capacity=32))
    tickets = [[None] * 10 for _ in shapes]

    def submit(k):
        for i in range(10):
            for s in range(len(shapes)):
                if (i + s) % 2 == k:
                    tickets[s][i] = engs[s].submit(qs[s][i])

    th = [threading.Thread(target=submit, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    got = [[None] * 10 for _ in shapes]
    for s in range(len(shapes)):
        for i in range(10):
            got[s][i] = engs[s].wait(tickets[s][i])[0]
            assert got[s][i] == solo[s][i], (s, i)
    for e in engs:
        e.close()
    for srv in srvs:
        srv.close()
    # one answer per database against the oracle
    for s, (rows, inst, d, seed) in enumerate(shapes):
        state = O.synth_state(seed, rows, inst)
        i = 3 + 4 * s
        parts = O.synth_query(seed * 10 + i, rows, nq=2)
        exp = O.respond_raw(state[0], rows, inst, d, *state[1:], *parts)
        assert got[s][i] == O.serialize_response(exp, []), (s, i)


# --- DB encode (encode_kernel.cu:18-228) -------------------------------------

@pytest.mark.parametrize("db_rows,d,groups", [(64, 1, 2), (64, 2, 3), (128, 8, 2), (64, 32, 2)])
def test_encode_matches_oracle(torch, db_rows, d, groups):
    say("encode", db_rows, d, groups)
    db_cols = groups * d * N
    rng = np.random.default_rng(db_rows + d)
    raw = rng.integers(0, 65536, size=db_rows * db_cols, dtype=np.uint16)
    raw[::131] = 65535  # non-canonical words (0 mod p)
    exp = O.encode(raw, db_rows, db_cols, d)
    dr = dev_from(torch, raw)
    out = empty(torch, raw.size, np.uint16)
    IA.encode_db_device(out, dr, db_rows, db_cols, d)
    assert (host(torch, out, np.uint16) == exp).all()
    IA.encode_db_device(dr, dr, db_rows, db_cols, d)  # in place
    assert (host(torch, dr, np.uint16) == exp).all()


def test_encode_rejects_bad_shapes(torch):
    say("encode bad shapes")
    dr = empty(torch, 64 * 4 * N, np.uint16)
    for rows, cols, d in [(64, 3 * N, 3), (60, 4 * N, 2), (64, 3 * N, 2), (64, 64 * N, 64)]:
        with pytest.raises(IA.InspireError):
            IA.encode_db_device(dr, dr, rows, cols, d)


def test_server_encode_db_equals_set_db_of_oracle_encoding(torch):
    say("server encode_db 2048x8 d4")
    db_rows, inst, d = 2048, 8, 4
    st = O.synth_state(23, db_rows, inst)
    rng = np.random.default_rng(29)
    raw = rng.integers(0, 65536, size=db_rows * inst * N, dtype=np.uint16)
    enc = O.encode(raw, db_rows, inst * N, d)
    y, z, rows, bodies = O.synth_query(24, db_rows, nq=2)
    q = O.serialize_query(y, z, rows, bodies)
    out = []
    for how in ("encode", "set"):
        srv = IA.Server(db_rows, inst, d)
        if how == "encode":
            srv.encode_db(raw)
        else:
            srv.set_db(enc)
        srv.set_precomp(st[1], st[2], st[3], st[4])
        out.append(srv.respond_serialized(q))
        srv.close()
    assert out[0] == out[1]
    exp = O.respond_raw(enc, db_rows, inst, d, st[1], st[2], st[3], st[4], y, z, rows, bodies)
    assert out[0] == O.serialize_response(exp, [])


def test_encode_full_size_column_group_matches_oracle(torch):
    # The 1M-key shape: db_rows 2048, 720,896 columns, d = 32 (11 column
    # groups); group 7 is checked against the oracle.
    say("encode 1M shape")
    db_rows, d, c = 2048, 32, 11
    db_cols = c * d * N
    raw = empty(torch, db_rows * db_cols, np.uint16)
    IA.synth_fill(raw, 19, IA.STREAM_DB, db_rows * db_cols, IA.KIND_U16)
    out = empty(torch, db_rows * db_cols, np.uint16)
    IA.encode_db_device(out, raw, db_rows, db_cols, d)
    wp = 7
    lo, hi = wp * d * N * db_rows, (wp + 1) * d * N * db_rows
    slab = host(torch, raw[lo:hi], np.uint16)
    exp = O.encode(slab, db_rows, d * N, d)
    assert (host(torch, out[lo:hi], np.uint16) == exp).all()
    del raw, out
    torch.cuda.empty_cache()


def test_legacy_encode_then_gemv(torch):
    # keyword.rs:404-424: gpu_encode_db_from_host, then the GEMV on the device
    # DB it keeps (gpu_encode_get_device_db); the fold slot and swap.
    say("legacy encode + gemv")
    import ctypes
    L = IA.lib()
    L.gpu_encode_get_device_db.restype = ctypes.c_void_p
    L.gpu_encode_get_device_db_fold.restype = ctypes.c_void_p
    db_rows, d, c = 2048, 4, 2
    db_cols = c * d * N
    rng = np.random.default_rng(41)
    raws = [rng.integers(0, 65536, size=db_rows * db_cols, dtype=np.uint16) for _ in range(2)]
    a = O.synth(83, O.STREAM_ROW, db_rows, O.KIND_CONDENSED)
    k = O.constants()
    u = ctypes.c_uint64
    vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)

    def gemv(ptr):
        out = np.zeros(db_cols, dtype=np.uint64)
        L.gpu_gemv_query(vp(out), vp(a), ctypes.c_void_p(ptr), db_rows, db_cols, u(O.Q0), u(O.Q1),
                         u(k["cr1_q0"]), u(k["cr1_q1"]), u(k["mod0_inv_mod1"]), u(k["mod1_inv_mod0"]), u(O.Q),
                         u(k["cr0_q"]), u(k["cr1_q"]))
        return out

    num_inv = pow(d, -1, 65535)
    L.gpu_encode_db_from_host(vp(raws[0]), None, db_rows, db_cols, d, N, c, ctypes.c_uint32(65535),
                              ctypes.c_uint32(num_inv))
    p0 = L.gpu_encode_get_device_db()
    assert p0
    exp0 = O.gemv(a, O.encode(raws[0], db_rows, db_cols, d), db_rows, db_cols)
    assert (gemv(p0) == exp0).all()
    L.gpu_encode_db_from_host_fold(vp(raws[1]), db_rows, db_cols, d, N, c, ctypes.c_uint32(65535),
                                   ctypes.c_uint32(num_inv))
    p1 = L.gpu_encode_get_device_db_fold()
    assert p1 and L.gpu_encode_get_device_db() == p0
    L.gpu_encode_swap_primary()
    assert L.gpu_encode_get_device_db() == p1 and not L.gpu_encode_get_device_db_fold()
    assert (gemv(p1) == O.gemv(a, O.encode(raws[1], db_rows, db_cols, d), db_rows, db_cols)).all()
    # a wrong num_inv is refused, leaving the primary in place
    L.gpu_encode_db_from_host(vp(raws[0]), None, db_rows, db_cols, d, N, c, ctypes.c_uint32(65535),
                              ctypes.c_uint32(num_inv + 1))
    assert L.gpu_encode_get_device_db() == p1
    L.gpu_encode_free_device_db()
    assert not L.gpu_encode_get_device_db()
    L.gpu_trim_memory_pool()


def test_legacy_adopt_path_matches_oracle(torch):
    # The calls KeywordServer::setup and the offline precomputation make
    # through the reference's FFI (keyword.rs:404-457, engine.rs:936-939): the
    # encoded device DB adopted by gpu_gemv_set_device_db and queried with
    # gpu_gemv_query; device-resident precomputation adopted by
    # gpu_packing_online_adopt_precomp, keys expanded on the device or uploaded
    # as host rotations (gpu_packing_online_upload_keys), and a second
    # precomputation swapped in (gpu_packing_online_swap_precomp).
    say("legacy adopt path")
    import ctypes
    L = IA.lib()
    L.gpu_encode_get_device_db.restype = ctypes.c_void_p
    vp = ctypes.c_void_p
    u = ctypes.c_uint64
    k = O.constants()
    ap = lambda x: x.ctypes.data_as(vp)
    L.gpu_init_memory_pool()
    db_rows, d, c = 2048, 2, 2
    inst = c * d
    db_cols = inst * N
    raw = np.random.default_rng(43).integers(0, 65536, size=db_rows * db_cols, dtype=np.uint16)
    L.gpu_encode_db_from_host(ap(raw), None, db_rows, db_cols, d, N, c, ctypes.c_uint32(65535),
                              ctypes.c_uint32(pow(d, -1, 65535)))
    enc = L.gpu_encode_get_device_db()
    assert enc
    d_db = vp()
    L.gpu_gemv_set_device_db(ctypes.byref(d_db), vp(enc), ctypes.c_size_t(raw.nbytes))
    assert d_db.value == enc
    a = O.synth(85, O.STREAM_ROW, db_rows, O.KIND_CONDENSED)
    got = np.zeros(db_cols, dtype=np.uint64)
    L.gpu_gemv_query(ap(got), ap(a), d_db, db_rows, db_cols, u(O.Q0), u(O.Q1), u(k["cr1_q0"]), u(k["cr1_q1"]),
                     u(k["mod0_inv_mod1"]), u(k["mod1_inv_mod0"]), u(O.Q), u(k["cr0_q"]), u(k["cr1_q"]))
    assert (got == O.gemv(a, O.encode(raw, db_rows, db_cols, d), db_rows, db_cols)).all()
    L.gpu_encode_free_device_db()

    tables = np.concatenate([O.automorph_table(t) for t in range(1, 2 * N, 2)]).astype(np.uint32)
    gen_pows = np.array([pow(5, i, 2 * N) for i in range(O.NROT)], dtype=np.uint32)
    L.gpu_rotation_upload_tables(ap(tables), ap(gen_pows), N, O.NROT, N)
    L.gpu_packing_online_ensure_key_buffers(O.NINNER, 3, N)
    y, z, _, _ = O.synth_query(86, 2048, nq=1)
    states = []
    for seed in (87, 88):
        host_pre = [O.synth(seed, O.STREAM_BT, inst * O.NINNER * N, O.KIND_CONDENSED),
                    O.synth(seed, O.STREAM_BTB, inst * O.NINNER * N, O.KIND_CONDENSED),
                    O.synth(seed, O.STREAM_BTH, inst * 3 * N, O.KIND_CONDENSED)]
        states.append((host_pre, [dev_from(torch, x) for x in host_pre]))

    def compute():
        out = np.zeros(inst * 2 * N, dtype=np.uint64)
        L.gpu_packing_online_compute(ap(out), inst, O.NROT, 3, N, 128, u(O.Q0), u(O.Q1), u(k["cr1_q0"]),
                                     u(k["cr1_q1"]))
        return out.reshape(inst, 2, N)

    (h0, d0), (h1, d1) = states
    L.gpu_packing_online_adopt_precomp(*[vp(t.data_ptr()) for t in d0], inst, O.NINNER, 3, N)
    L.gpu_rotation_expand_keys(vp(L.gpu_packing_online_get_d_y_all()), vp(L.gpu_packing_online_get_d_y_bar_all()),
                               ap(y), O.NROT, 3, N)
    L.gpu_packing_online_upload_z_body(ap(z), 3, N)
    exp0 = O.packing(*h0, y, z, inst)
    assert (compute() == exp0).all()
    # the same keys as host rotations: y_all[i][k] = y[k] through the table of
    # g_i = 5^i, y_bar_all[i][k] through the table of 2n - g_i (rotation_kernel.cu:21-59)
    tab = tables.reshape(N, N)
    y_all = np.stack([y[:, tab[(g - 1) // 2]] for g in gen_pows])
    y_bar_all = np.stack([y[:, tab[(2 * N - g - 1) // 2]] for g in gen_pows])
    L.gpu_packing_online_upload_keys(ap(np.ascontiguousarray(y_all)), ap(np.ascontiguousarray(y_bar_all)), ap(z),
                                     O.NINNER, 3, N)
    assert (compute() == exp0).all()
    L.gpu_packing_online_swap_precomp(*[vp(t.data_ptr()) for t in d1], inst, O.NINNER, 3, N)
    assert (compute() == O.packing(*h1, y, z, inst)).all()
    L.gpu_packing_online_free()
    L.gpu_rotation_free()
    torch.cuda.synchronize()
    del states, d0, d1


def test_engine_live_update_swaps_state(torch):
    # Live DB update: a staged server of the same shape with another DB,
    # precomputation and stash replaces the engine's state between batches.
    # Every answer equals the same query answered alone on the state it ran
    # on, stash included, and never mixes the two.
    say("engine live update")
    import time as _time
    db_rows, inst, d = 2048, 8, 4
    live = IA.Server(db_rows, inst, d)
    live.fill_synthetic(61)
    live.set_stash([b"A" * 16])
    staged = IA.Server(db_rows, inst, d)
    staged.fill_synthetic(62)
    staged.set_stash([b"B" * 16, b"C" * 16])
    qs = [q for q, _ in _queries(db_rows, [700 + i for i in range(6)])]
    solo_a = [live.respond_serialized(q) for q in qs]
    solo_b = [staged.respond_serialized(q) for q in qs]
    assert all(a != b for a, b in zip(solo_a, solo_b))
    eng = IA.Engine(live, nq=2, max_batch=4, window_ms=0.0, capacity=32)
    # 1: answered on the old state before the swap
    t = [eng.submit(q) for q in qs[:2]]
    assert [eng.wait(x)[0] for x in t] == solo_a[:2]
    # 2: dispatched before the swap, collected after it: old state, old stash
    t = eng.submit(qs[2])
    for _ in range(2000):
        if eng.stats()["queries"] >= 3:
            break
        _time.sleep(0.001)
    assert eng.stats()["queries"] >= 3
    eng.swap_state(staged)
    assert eng.wait(t)[0] == solo_a[2]
    # 3: after the swap, the new state and its stash
    t = [eng.submit(q) for q in qs[3:]]
    assert [eng.wait(x)[0] for x in t] == solo_b[3:]
    # the staged handle now holds the old state
    assert staged.respond_serialized(qs[0]) == solo_a[0]
    # and back again while queries are queued: each is wholly one state or the other
    t = [eng.submit(q) for q in qs]
    eng.swap_state(staged)
    for i, x in enumerate(t):
        r = eng.wait(x)[0]
        assert r in (solo_a[i], solo_b[i]), i
    t = [eng.submit(q) for q in qs]
    assert [eng.wait(x)[0] for x in t] == solo_a
    with pytest.raises(IA.InspireError):
        live.swap_state(IA.Server(db_rows, 4, d))
    eng.close()
    live.close()
    staged.close()
