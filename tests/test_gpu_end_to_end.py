"""End-to-end on the GPU: the DB is encoded on the device, the server holds real
offline state (hint, prep-pack and InspiRING precompute of that DB, built by
the oracle), clients send real encrypted queries, and each client decrypts the
GPU's response to the stored values of the selected row and sub-column
(tests/e2e_mlkem.rs:27-38 at the level of the answer path).  The shape is the
10,000-key one: db_rows 2048, 8 instances, interpolation degree 8 (one chain).
Responses are also compared byte for byte with the oracle's."""
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

pytestmark = pytest.mark.gpu
N = O.N
DB_ROWS, INST, D = 2048, 8, 8


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    return _t


@pytest.fixture(scope="module")
def real():
    say("end-to-end state: encode + offline precompute, 2048 x 8 instances, d 8")
    rng = np.random.default_rng(77)
    raw = rng.integers(0, 65536, size=INST * N * DB_ROWS, dtype=np.uint16)
    raw.reshape(INST * N, DB_ROWS)[3:40, 100] = 0xFFFF
    enc = O.encode(raw, DB_ROWS, INST * N, D)
    st = O.setup_state(enc, DB_ROWS, INST)
    srv = IA.Server(DB_ROWS, INST, D)
    srv.encode_db(raw)
    srv.set_precomp(*st)
    yield raw.reshape(INST * N, DB_ROWS), enc, st, srv
    srv.close()


def split_response(resp):
    """wire.rs:708-726 -> list of per-hash-query blobs."""
    n = int.from_bytes(resp[:4], "little")
    out, o = [], 4
    for _ in range(n):
        ln = int.from_bytes(resp[o:o + 4], "little")
        out.append(resp[o + 4:o + 4 + ln])
        o += 4 + ln
    return out


def expected(rawm, row, sub):
    c = INST // D
    return np.stack([rawm[(w * D + sub) * N:(w * D + sub + 1) * N, row] for w in range(c)]).astype(np.uint64) % 65535


def client(seed, hamming, targets):
    sk = O.client_secret(seed, hamming)
    y, z, rows, bodies = O.client_query(sk, DB_ROWS, D, targets, seed=seed + 1000)
    return sk, (y, z, rows, bodies), O.serialize_query(y, z, rows, bodies)


@pytest.mark.parametrize("packing", ["mfma", "valu"])
@pytest.mark.parametrize("hamming", [0, O.HAMMING_WEIGHT])
def test_gpu_respond_decrypts_to_stored_values(torch, real, packing, hamming):
    rawm, enc, st, srv = real
    say("e2e respond", packing, hamming)
    srv.set_packing_kernel(IA.PACKING_MFMA if packing == "mfma" else IA.PACKING_VALU)
    targets = [(100, 0), (2047, 5)]
    sk, parts, q = client(31 + hamming, hamming, targets)
    resp = srv.respond_serialized(q)
    exp_bytes = O.serialize_response(O.respond_raw(enc, DB_ROWS, INST, D, *st, *parts), [])
    assert resp == exp_bytes
    for (row, sub), blob in zip(targets, split_response(resp)):
        assert np.array_equal(O.decrypt_response(blob, sk), expected(rawm, row, sub)), (row, sub)
    assert (expected(rawm, 100, 0)[0][3:40] == 0).all()
    srv.set_packing_kernel(IA.PACKING_MFMA)


@pytest.mark.parametrize("count", [5, 20, 70])
def test_gpu_batched_clients_decrypt_their_own_values(torch, real, count):
    # Clients with their own keys and targets in one batched respond: each one
    # decrypts its own row and sub-column (packing GEMM over all clients' keys,
    # batched GEMV and fold).
    rawm, enc, st, srv = real
    say("e2e batch", count)
    srv.set_max_batch(max(32, count))
    rng = np.random.default_rng(count)
    clients = []
    for i in range(count):
        targets = [(int(rng.integers(0, DB_ROWS)), int(rng.integers(0, D))) for _ in range(2)]
        clients.append((targets,) + client(500 + i, O.HAMMING_WEIGHT if i % 3 else 0, targets))
    qlen = len(clients[0][3])
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * count, dtype=np.uint8)
    for i, cl in enumerate(clients):
        host[i * stride:i * stride + qlen] = np.frombuffer(cl[3], dtype=np.uint8)
    dq = torch.from_numpy(host).cuda()
    blob = 2 * (INST // D) * O.RESP_BYTES
    dout = torch.zeros(count * blob, dtype=torch.uint8, device="cuda")
    srv.respond_batch_device(dq, stride, qlen, 2, count, dout, blob)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    for i, (targets, sk, parts, q) in enumerate(clients):
        mine = out[i * blob:(i + 1) * blob]
        for h, (row, sub) in enumerate(targets):
            part = mine[h * blob // 2:(h + 1) * blob // 2]
            assert np.array_equal(O.decrypt_response(part, sk), expected(rawm, row, sub)), (i, h)
    # the first client's bytes also against the oracle
    exp = O.respond_raw(enc, DB_ROWS, INST, D, *st, *clients[0][2])
    assert out[:blob] == b"".join(exp)


@pytest.mark.parametrize("db_rows,inst", [(DB_ROWS, INST), (4096, 2)])
def test_gpu_offline_setup_matches_oracle(torch, real, db_rows, inst):
    # hint + prep-pack + InspiRING precompute + collapse on the device equal the
    # oracle's restatement of the reference's CPU path, word for word.
    say("setup_offline_device", db_rows, inst)
    if db_rows == DB_ROWS and inst == INST:
        _, enc, st, _ = real
    else:
        rng = np.random.default_rng(db_rows + inst)
        enc = rng.integers(0, 65535, size=inst * N * db_rows, dtype=np.uint16)
        st = O.setup_state(enc, db_rows, inst)
    d_db = torch.from_numpy(enc.view(np.int16)).cuda()
    outs = [torch.zeros(a.size, dtype=torch.int64, device="cuda") for a in st]
    IA.setup_offline_device(d_db, db_rows, inst, *outs)
    torch.cuda.synchronize()
    for name, got, exp in zip(("bold_t", "bold_t_bar", "bold_t_hat", "a_hat"), outs, st):
        g = got.cpu().numpy().view(np.uint64)
        assert np.array_equal(g, exp), (name, int(np.argmax(g != exp)))


def test_gpu_encode_setup_respond_decrypts(torch, real):
    # The whole server side on the device: encode, offline precomputation,
    # respond; the client decrypts the stored values.
    rawm, enc, st, _ = real
    say("encode + setup_offline + respond")
    srv = IA.Server(DB_ROWS, INST, D)
    srv.encode_db(np.ascontiguousarray(rawm).reshape(-1))
    srv.setup_offline()
    targets = [(5, 7), (1500, 2)]
    sk, parts, q = client(901, O.HAMMING_WEIGHT, targets)
    resp = srv.respond_serialized(q)
    assert resp == O.serialize_response(O.respond_raw(enc, DB_ROWS, INST, D, *st, *parts), [])
    for (row, sub), blob in zip(targets, split_response(resp)):
        assert np.array_equal(O.decrypt_response(blob, sk), expected(rawm, row, sub)), (row, sub)
    srv.close()


def test_lean_server_matches_full_server(torch, real):
    # A lean server (fragment-ordered bold_t only) answers with the same bytes
    # as the full server: from synthetic state, from set_precomp and from the
    # device setup; it refuses the VALU packing.
    rawm, enc, st, srv = real
    say("lean server")
    lean = IA.Server(DB_ROWS, INST, D, lean=True)
    full = IA.Server(DB_ROWS, INST, D)
    with pytest.raises(IA.InspireError):
        lean.set_packing_kernel(IA.PACKING_VALU)
    sk, parts, q = client(77, O.HAMMING_WEIGHT, [(9, 3), (700, 6)])
    # synthetic state
    lean.fill_synthetic(5)
    full.fill_synthetic(5)
    full.set_packing_kernel(IA.PACKING_MFMA)
    assert lean.respond_serialized(q) == full.respond_serialized(q)
    # real state through set_precomp, then through the device setup
    lean.set_db(enc)
    lean.set_precomp(*st)
    exp = srv.respond_serialized(q)
    assert lean.respond_serialized(q) == exp
    lean.encode_db(np.ascontiguousarray(rawm).reshape(-1))
    lean.setup_offline()
    assert lean.respond_serialized(q) == exp
    # batched on the lean server
    lean.set_max_batch(8)
    qlen = len(q)
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * 3, dtype=np.uint8)
    for i in range(3):
        host[i * stride:i * stride + qlen] = np.frombuffer(q, dtype=np.uint8)
    blob = 2 * (INST // D) * O.RESP_BYTES
    dout = torch.zeros(3 * blob, dtype=torch.uint8, device="cuda")
    lean.respond_batch_device(torch.from_numpy(host).cuda(), stride, qlen, 2, 3, dout, blob)
    torch.cuda.synchronize()
    out = dout.cpu().numpy().tobytes()
    body = split_response(exp)
    for i in range(3):
        assert out[i * blob:(i + 1) * blob] == body[0] + body[1]
    lean.close()
    full.close()


def test_dim0_65536_shape_decrypts(torch):
    # The 32M-key serving shape has db_rows = dim0 = 65536 (32 query RLWEs per
    # hash query, a 65536-term GEMV): device encode + setup + respond at that
    # row count with 2 instances, against the oracle and through decryption.
    rows, inst, d = 65536, 2, 2
    say("dim0 65536: 65536 x 2 instances, d 2")
    rng = np.random.default_rng(65536)
    raw = rng.integers(0, 65535, size=inst * N * rows, dtype=np.uint16)
    enc = O.encode(raw, rows, inst * N, d)
    st = O.setup_state(enc, rows, inst)
    srv = IA.Server(rows, inst, d, lean=True)
    srv.encode_db(raw)
    srv.setup_offline()
    sk = O.client_secret(3, O.HAMMING_WEIGHT)
    targets = [(65535, 1), (12345, 0)]
    y, z, qrows, bodies = O.client_query(sk, rows, d, targets, seed=4)
    q = O.serialize_query(y, z, qrows, bodies)
    resp = srv.respond_serialized(q)
    assert resp == O.serialize_response(O.respond_raw(enc, rows, inst, d, *st, y, z, qrows, bodies), [])
    rawm = raw.reshape(inst * N, rows)
    for (row, sub), blob in zip(targets, split_response(resp)):
        exp = np.stack([rawm[(w * d + sub) * N:(w * d + sub + 1) * N, row] for w in range(inst // d)])
        assert np.array_equal(O.decrypt_response(blob, sk), exp.astype(np.uint64)), (row, sub)
    srv.close()
