"""Parity at the benchmark's full size: 1,000,000 keys at the reference's
default shape (db_rows 2048, 720,896 columns, 352 packed outputs, d 32, 11
fold chains), synthetic device-resident state.  The oracle recomputes whole
fold chains from slices of the same synthetic state (the counter-based
generator gives any slice directly): the 32 packed outputs of a chain, the
GEMV over its 65,536 columns, the fold and the modulus switch.  Each chain's
12,288 response bytes must be identical to the GPU's: every chain of a
single query, queries in batches of 70, 128 and 256 (every query also against
its own single-query answer), 1,000 sequential single-query responds against
the batched path with a sample against the oracle, and the offline setup of
two outputs."""
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O
from chain_oracle import chain_blobs, chain_of, split_response, stage_queries

pytestmark = pytest.mark.gpu
N = O.N
ROWS, INST, D = 2048, 352, 32
C = INST // D
SEED = 11
BLOB = 2 * C * O.RESP_BYTES


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
    srv.set_max_batch(256)
    yield srv
    srv.close()
    torch.cuda.empty_cache()


def oracle_chains(w, q):
    return chain_blobs(SEED, ROWS, D, w, *q)


@pytest.mark.parametrize("small", [False, True])
def test_full_size_single_query_chains_match_oracle(server, small):
    # uniform residues (a client with a real ternary key): all 11 chains;
    # small residues (the reference's sk = 0 client): the first and last
    chains = [0, 10] if small else list(range(C))
    say("1M single query, chains", chains, "small =", small)
    q = O.synth_query(13 + small, ROWS, nq=2, small=small)
    blobs = split_response(server.respond_serialized(O.serialize_query(*q)))
    assert len(blobs) == 2 and all(len(b) == C * O.RESP_BYTES for b in blobs)
    for w in chains:
        exp = oracle_chains(w, q)
        for h in range(2):
            assert chain_of(blobs[h], w) == exp[h], (w, h)


_QUERIES = {}


def _batch_queries(count):
    """count synthetic keyword queries and their single-query answers (cached
    across the batch sizes)."""
    qs = _QUERIES.setdefault("q", [])
    while len(qs) < count:
        qs.append(O.synth_query(500 + len(qs), ROWS, nq=2))
    return qs[:count], [O.serialize_query(*q) for q in qs[:count]]


@pytest.mark.parametrize("count,checks", [(70, [(67, 5)]), (128, [(0, 0), (127, 10)]),
                                          (256, [(64, 3), (131, 7), (255, 10)])])
def test_full_size_batched_chains_match_oracle(server, count, checks):
    # groups of 64 clients per packing pass: 70 = one full and one partial
    # group, 128 and 256 two and four full groups.  Every query equals its
    # single-query answer; (query, chain) pairs in `checks` equal the oracle.
    import torch
    say("1M batch of %d, oracle checks %s" % (count, checks))
    qs, ser = _batch_queries(count)
    dq, qlen, stride = stage_queries(torch, ser)
    dout = torch.zeros(count * BLOB, dtype=torch.uint8, device="cuda")
    server.respond_batch_device(dq, stride, qlen, 2, count, dout, BLOB)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    solo = _QUERIES.setdefault("solo", {})
    for i in range(count):
        if i not in solo:
            s = split_response(server.respond_serialized(ser[i]))
            solo[i] = s[0] + s[1]
        assert out[i * BLOB:(i + 1) * BLOB] == solo[i], i
    half = BLOB // 2
    for i, w in checks:
        exp = oracle_chains(w, qs[i])
        mine = out[i * BLOB:(i + 1) * BLOB]
        for h in range(2):
            assert chain_of(mine[h * half:(h + 1) * half], w) == exp[h], (i, w, h)


def test_full_size_sequential_queries_match_batched_and_oracle(server):
    # 1,000 sequential single-query responds (int8 MFMA GEMV and packing, the
    # 1024-thread fold) against the same queries answered in batches of 250
    # (the LDS-ring GEMV, the 64-client packing, the sequential-prime fold):
    # every response identical; eight of them, spread over the chains, also
    # equal the oracle.
    import torch
    n, B = 1000, 250
    say("1M: %d sequential queries vs batches of %d" % (n, B))
    qs = [O.synth_query(9000 + i, ROWS, nq=2, small=(i % 5 == 4)) for i in range(n)]
    ser = [O.serialize_query(*q) for q in qs]
    dq, qlen, stride = stage_queries(torch, ser)
    single = torch.zeros(n * BLOB, dtype=torch.uint8, device="cuda")
    st = torch.cuda.Stream()
    for i in range(n):
        server.respond_device(dq.data_ptr() + i * stride, qlen, 2, single.data_ptr() + i * BLOB, st.cuda_stream)
    batched = torch.zeros(n * BLOB, dtype=torch.uint8, device="cuda")
    for b in range(0, n, B):
        server.respond_batch_device(dq.data_ptr() + b * stride, stride, qlen, 2, B, batched.data_ptr() + b * BLOB,
                                    BLOB, st.cuda_stream)
    torch.cuda.synchronize()
    s_host = single.cpu().numpy()
    b_host = batched.cpu().numpy()
    for i in range(n):
        assert np.array_equal(s_host[i * BLOB:(i + 1) * BLOB], b_host[i * BLOB:(i + 1) * BLOB]), i
    half = BLOB // 2
    for k, i in enumerate((0, 141, 287, 404, 519, 650, 777, 999)):
        w = k * 3 % C
        exp = oracle_chains(w, qs[i])
        mine = s_host[i * BLOB:(i + 1) * BLOB].tobytes()
        for h in range(2):
            assert chain_of(mine[h * half:(h + 1) * half], w) == exp[h], (i, w, h)


def test_offline_setup_in_two_passes_matches_oracle():
    # One pass holds as many instances as keep all their collapse chains
    # resident (384 on 256 CUs), so 400 instances run in two passes; the first
    # and last output of each pass against the oracle.
    import torch
    inst = 400
    say("offline setup, %d outputs in two passes" % inst)
    ncell = ROWS * inst * N
    d_db = torch.empty(ncell, dtype=torch.int16, device="cuda")
    IA.synth_fill(d_db, 29, O.STREAM_DB, ncell, O.KIND_U16)
    outs = [torch.zeros(n, dtype=torch.int64, device="cuda")
            for n in (inst * O.NINNER * N, inst * O.NINNER * N, inst * 3 * N, inst * N)]
    IA.setup_offline_device(d_db, ROWS, inst, *outs)
    torch.cuda.synchronize()
    per = (O.NINNER * N, O.NINNER * N, 3 * N, N)
    for t in (0, 383, 384, inst - 1):
        db = O.synth(29, O.STREAM_DB, N * ROWS, O.KIND_U16, start=t * N * ROWS)
        exp = O.setup_precomp(O.setup_atilde(db, ROWS, 0, N), 1)
        for name, got, e, n in zip(("bold_t", "bold_t_bar", "bold_t_hat", "a_hat"), outs, exp, per):
            g = got[t * n:(t + 1) * n].cpu().numpy().view(np.uint64)
            assert np.array_equal(g, e), (t, name)
    del d_db, outs
    torch.cuda.empty_cache()


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
    del d_db, outs
    torch.cuda.empty_cache()
