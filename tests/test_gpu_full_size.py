"""Parity at the benchmark's full size: 1,000,000 keys at the reference's
default shape (db_rows 2048, 720,896 columns, 352 packed outputs, d 32, 11
fold chains), synthetic device-resident state.  The oracle recomputes whole
fold chains from slices of the same synthetic state (the counter-based
generator gives any slice directly): the 32 packed outputs of a chain, the
GEMV over its 65,536 columns, the fold and the modulus switch.  Each chain's
12,288 response bytes must be identical to the GPU's, for the single-query
path, for a query inside a batch of 70, and for the offline setup of two
outputs."""
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

pytestmark = pytest.mark.gpu
N = O.N
ROWS, INST, D = 2048, 352, 32
SEED = 11


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def server():
    import torch
    assert torch.cuda.is_available()
    say("1M-key state: fill_synthetic(%d), 352 instances" % SEED)
    srv = IA.Server(ROWS, INST, D)
    srv.fill_synthetic(SEED)
    srv.set_packing_kernel(IA.PACKING_MFMA)
    yield srv
    srv.close()


def chain_blobs(w, y, z, rows, bodies):
    """Oracle: chain w's response blob for each hash query."""
    j0 = w * D
    bt = O.synth(SEED, O.STREAM_BT, D * O.NINNER * N, O.KIND_CONDENSED, start=j0 * O.NINNER * N)
    btb = O.synth(SEED, O.STREAM_BTB, D * O.NINNER * N, O.KIND_CONDENSED, start=j0 * O.NINNER * N)
    bth = O.synth(SEED, O.STREAM_BTH, D * 3 * N, O.KIND_CONDENSED, start=j0 * 3 * N)
    ahat = O.synth(SEED, O.STREAM_AHAT, D * N, O.KIND_MODQ, start=j0 * N).reshape(D, N)
    db = O.synth(SEED, O.STREAM_DB, D * N * ROWS, O.KIND_U16, start=j0 * N * ROWS)
    P = O.post(O.packing(bt, btb, bth, y, z, D))
    del bt, btb
    neg_a, _ = O.rgsw_neg_a_ntt()
    out = []
    for h in range(rows.shape[0]):
        c = O.gemv(rows[h], db, ROWS, D * N).reshape(D, N)
        ct = O.fold_chain(ahat, c + P, neg_a, bodies[h])
        out.append(O.switch_and_keep(ct))
    return out


def split_response(resp):
    n = int.from_bytes(resp[:4], "little")
    out, o = [], 4
    for _ in range(n):
        ln = int.from_bytes(resp[o:o + 4], "little")
        out.append(resp[o + 4:o + 4 + ln])
        o += 4 + ln
    return out


@pytest.mark.parametrize("small", [False, True])
def test_full_size_single_query_chains_match_oracle(server, small):
    say("1M single query, chains 0 and 10, small =", small)
    y, z, rows, bodies = O.synth_query(13 + small, ROWS, nq=2, small=small)
    resp = server.respond_serialized(O.serialize_query(y, z, rows, bodies))
    blobs = split_response(resp)
    assert len(blobs) == 2 and all(len(b) == 11 * O.RESP_BYTES for b in blobs)
    for w in (0, 10):
        exp = chain_blobs(w, y, z, rows, bodies)
        for h in range(2):
            assert blobs[h][w * O.RESP_BYTES:(w + 1) * O.RESP_BYTES] == exp[h], (w, h)


def test_full_size_batched_chain_matches_oracle(server):
    # 70 keyword queries (two 64-client packing groups); query 67 (second
    # group) checked on chain 5 against the oracle, and against its own
    # single-query response on every chain.
    import torch
    say("1M batch of 70, query 67, chain 5")
    count = 70
    qs = [O.synth_query(500 + i, ROWS, nq=2) for i in range(count)]
    ser = [O.serialize_query(*q) for q in qs]
    qlen = len(ser[0])
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * count, dtype=np.uint8)
    for i, b in enumerate(ser):
        host[i * stride:i * stride + qlen] = np.frombuffer(b, dtype=np.uint8)
    server.set_max_batch(count)
    blob = 2 * 11 * O.RESP_BYTES
    dout = torch.zeros(count * blob, dtype=torch.uint8, device="cuda")
    server.respond_batch_device(torch.from_numpy(host).cuda(), stride, qlen, 2, count, dout, blob)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    mine = out[67 * blob:68 * blob]
    solo = split_response(server.respond_serialized(ser[67]))
    assert mine == solo[0] + solo[1]
    exp = chain_blobs(5, *qs[67])
    half = blob // 2
    for h in range(2):
        assert mine[h * half + 5 * O.RESP_BYTES:h * half + 6 * O.RESP_BYTES] == exp[h], h


def test_full_size_offline_setup_matches_oracle():
    # The offline precomputation over the whole 1M-key DB (synthetic encoded
    # DB) on the device; outputs 0 and 351 against the oracle.
    import torch
    say("1M offline setup, outputs 0 and 351")
    ncell = ROWS * INST * N
    d_db = torch.empty(ncell, dtype=torch.int16, device="cuda")
    IA.synth_fill(d_db, 23, O.STREAM_DB, ncell, O.KIND_U16)
    outs = [torch.zeros(n, dtype=torch.int64, device="cuda")
            for n in (INST * O.NINNER * N, INST * O.NINNER * N, INST * 3 * N, INST * N)]
    IA.setup_offline_device(d_db, ROWS, INST, *outs)
    torch.cuda.synchronize()
    per = (O.NINNER * N, O.NINNER * N, 3 * N, N)
    for t in (0, INST - 1):
        db = O.synth(23, O.STREAM_DB, N * ROWS, O.KIND_U16, start=t * N * ROWS)
        exp = O.setup_precomp(O.setup_atilde(db, ROWS, 0, N), 1)
        for name, got, e, n in zip(("bold_t", "bold_t_bar", "bold_t_hat", "a_hat"), outs, exp, per):
            g = got[t * n:(t + 1) * n].cpu().numpy().view(np.uint64)
            assert np.array_equal(g, e), (t, name)
