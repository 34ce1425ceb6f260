"""KeywordClient in the library (inspire_client_*, src/pir/client/keyword.rs:45-311)
and the TCP framing (inspire_amd.wire, src/protocol/wire.rs:15-86), host only.

The client is checked against the oracle, not against itself:
- the oracle's CPU server (O.respond_raw) answers the product client's
  queries on a keyword database laid out as the cuckoo table lays it out
  (bucket b at row b / per, columns (b % per) * ise ..), and the client must
  extract every stored value, find stash entries and return None for absent
  keys (the reference's own end-to-end check, tests/e2e_mlkem.rs:27-38);
- under the client's secret, its packing keys are key-switching keys
  (body + s mask - tau(s) 2^(19k) is small noise, packing.rs:3272-3294), its
  query rows decrypt to floor(q/p) at the target row and 0 elsewhere
  (client/online.rs:226-357), and its RGSW bodies decrypt to sigma and
  s sigma (client/keyword.rs:150-178);
- the noise has the reference's spread (discrete_gaussian.rs dg_seems_okay:
  standard deviation 16.042421 / sqrt(2 pi) within 10%, mean within 5 of its
  standard errors).
"""
import json
import math
import socket
import threading

import numpy as np
import pytest

import inspire_amd as IA
from inspire_amd import wire
import oracle as O

N = O.N
RECORDS, KS, VS = 200_000, 32, 32


@pytest.fixture(scope="module")
def kv_state():
    plan = IA.plan_shape(RECORDS, KS, VS)
    db_rows, inst, d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
    assert (db_rows, inst, d) == (2048, 4, 4)
    pp = dict(num_items=plan["num_items"], item_size_bits=8 * (KS + VS), dim0=plan["dim0"],
              cuckoo_seed=O.CUCKOO_SEED, num_hashes=2, key_size=KS, value_size=VS, entry_size=KS + VS,
              num_records=RECORDS)
    per, ise = pp["num_items"] // db_rows, (KS + VS) // 2
    rng = np.random.default_rng(5)
    raw = rng.integers(0, 65535, size=(inst * N, db_rows), dtype=np.uint16)
    items = []
    for i in range(6):
        key = rng.integers(0, 256, KS, dtype=np.uint8).tobytes()
        val = rng.integers(0, 256, VS, dtype=np.uint8).tobytes()
        words = np.frombuffer(key + val, dtype=">u2")
        if (words == 0xFFFF).any():   # a stored 0xFFFF reads back as 0 (mod p = 65535)
            continue
        b = O.cuckoo_positions(key, pp["num_items"])[i % 2]
        row, col = b // per, (b % per) * ise
        raw[col:col + ise, row] = words
        items.append((key, val))
    enc = O.encode(raw.reshape(-1), db_rows, inst * N, d)
    st = O.setup_state(enc, db_rows, inst)
    return pp, (db_rows, inst, d), enc, st, items


def serve(state, query, stash=()):
    pp, (db_rows, inst, d), enc, st, _ = state
    y, z, rows, bodies = O.deserialize_query(query.data, db_rows)
    blobs = O.respond_raw(enc, db_rows, inst, d, *st, y, z, rows, bodies)
    return O.serialize_response(blobs, list(stash))


def centered(v, q):
    v = np.asarray(v, dtype=np.int64)
    return np.where(v > q // 2, v - q, v)


@pytest.mark.parametrize("hamming", [0, 32])
def test_client_extracts_values_the_oracle_serves(kv_state, hamming):
    pp, _, _, _, items = kv_state
    cl = IA.KeywordClient(pp, hamming=hamming, rng_seed=100 + hamming)
    sk = cl.secret()
    assert (sk == 1).sum() == hamming and (sk == O.Q - 1).sum() == hamming and (sk == 0).sum() == N - 2 * hamming
    for key, val in items[:3]:
        q = cl.query(key)
        assert q.positions == O.cuckoo_positions(key, pp["num_items"]) == cl.positions(key)
        assert cl.extract(key, q, serve(kv_state, q)) == val
    # a key in no bucket: None; the same key in the stash: its value
    absent = bytes(range(KS))
    q = cl.query(absent)
    stash_value = bytes(range(100, 100 + VS))
    assert cl.extract(absent, q, serve(kv_state, q)) is None
    assert cl.extract(absent, q, serve(kv_state, q, [absent + stash_value])) == stash_value
    # a key of the wrong length is never found (extract returns None)
    assert cl.extract(absent[:-1], q, serve(kv_state, q)) is None


def test_query_wire_layout_and_noise(kv_state):
    pp, (db_rows, inst, d), _, _, items = kv_state
    cl = IA.KeywordClient(pp, hamming=0, rng_seed=7)
    key = items[0][0]
    q = cl.query(key)
    # 1 header byte + (2 x 3 x 2048 key values + 2 x (db_rows + 6 x 2048)) x 56 bits
    assert len(q.data) == cl.query_bytes() == 286_721 and q.data[0] == 2
    y, z, rows, bodies = O.deserialize_query(q.data, db_rows)
    # the product serializer and the client agree with the oracle's serializer
    assert O.serialize_query(y, z, rows, bodies) == q.data == IA.serialize_query(y, z, rows, bodies)
    # zero secret: each query row value is e (+ floor(q/p) at the target row), the same integer in both residues
    per = pp["num_items"] // db_rows
    noise = []
    for h, b in enumerate(q.positions):
        e0 = centered(rows[h] & np.uint64(0xFFFFFFFF), O.Q0)
        e1 = centered(rows[h] >> np.uint64(32), O.Q1)
        target = b // per
        scale = O.Q // 65535
        e0[target] = centered((int(rows[h][target] & 0xFFFFFFFF) - scale) % O.Q0, O.Q0)
        e1[target] = centered((int(rows[h][target] >> 32) - scale) % O.Q1, O.Q1)
        assert np.array_equal(e0, e1) and np.abs(e0).max() <= 65
        noise.append(e0)
    e = np.concatenate(noise).astype(np.float64)
    sd = 16.042421 / math.sqrt(2 * math.pi)
    assert abs(e.mean()) < 5 * sd / math.sqrt(len(e))
    assert abs(e.std() - sd) < 0.1 * sd


def test_packing_keys_and_rgsw_bodies_decrypt(kv_state):
    pp, (db_rows, inst, d), _, _, items = kv_state
    cl = IA.KeywordClient(pp, hamming=32, rng_seed=11)
    sk = cl.secret()
    key = items[1][0]
    q = cl.query(key)
    y, z, rows, bodies = O.deserialize_query(q.data, db_rows)
    s_ntt = O.ntt_forward(np.stack([sk % O.Q0, sk % O.Q1])[None])[0]

    def small(res):   # [2][N] residues -> the centered raw polynomial, checked equal in both primes
        r = O.ntt_inverse(np.stack(res)[None])[0]
        a, b = centered(r[0], O.Q0), centered(r[1], O.Q1)
        assert np.array_equal(a, b)
        return a

    w_seed, v_seed = [7] * 32, [8, 8] + [7] * 30
    for body, seed, t in ((y, w_seed, 5), (z, v_seed, 2 * N - 1)):
        mask = O.chacha20_u64(seed, 3 * 2 * N).reshape(3, 2, N)
        mask[:, 0] %= O.Q0
        mask[:, 1] %= O.Q1
        tau = O.automorph_raw(sk, t, O.Q)
        t_ntt = O.ntt_forward(np.stack([tau % O.Q0, tau % O.Q1])[None])[0]
        for k in range(3):
            res = []
            for c, qc in enumerate((O.Q0, O.Q1)):
                b = (body[k] >> np.uint64(32 * c)) & np.uint64(0xFFFFFFFF)
                res.append(np.array([(int(b[i]) + int(s_ntt[c][i]) * int(mask[k, c, i])
                                      - int(t_ntt[c][i]) * (1 << (19 * k))) % qc for i in range(N)], dtype=np.uint64))
            assert np.abs(small(res)).max() <= 65, (t, k)
    # RGSW columns: body - s a = e + msg, msg = sigma (column 2j+1) or s sigma (column 2j)
    a = O.chacha20_u64([7] * 32, N) % np.uint64(O.Q)
    a_ntt = O.ntt_forward(np.stack([a % O.Q0, a % O.Q1])[None])[0]
    per, ise = pp["num_items"] // db_rows, pp["item_size_bits"] // 16
    for h, bk in enumerate(q.positions):
        sub = (((bk % per) * ise) % (d * N)) // N
        ex = O.rgsw_exponent(sub, d)
        for j in range(3):
            sig = np.zeros(N, dtype=np.uint64)
            g = 1 << (19 * j)
            sig[ex % N] = g if ex < N else O.Q - g
            sig_ntt = O.ntt_forward(np.stack([sig % O.Q0, sig % O.Q1])[None])[0]
            for col, use_s in ((2 * j + 1, False), (2 * j, True)):
                res = []
                for c, qc in enumerate((O.Q0, O.Q1)):
                    msg = [(int(s_ntt[c][i]) * int(sig_ntt[c][i])) % qc if use_s else int(sig_ntt[c][i])
                           for i in range(N)]
                    res.append(np.array([(int(bodies[h][col][c][i]) - int(s_ntt[c][i]) * int(a_ntt[c][i]) - msg[i])
                                         % qc for i in range(N)], dtype=np.uint64))
                assert np.abs(small(res)).max() <= 65, (h, j, col)


def test_client_randomness(kv_state):
    pp, _, _, _, items = kv_state
    key = items[0][0]
    a = IA.KeywordClient(pp, hamming=16, rng_seed=3).query(key).data
    assert IA.KeywordClient(pp, hamming=16, rng_seed=3).query(key).data == a
    assert IA.KeywordClient(pp, hamming=16, rng_seed=4).query(key).data != a
    c1, c2 = IA.KeywordClient(pp, hamming=16), IA.KeywordClient(pp, hamming=16)
    assert not np.array_equal(c1.secret(), c2.secret())   # OS randomness
    # one client's successive queries use fresh noise
    cl = IA.KeywordClient(pp, hamming=0, rng_seed=5)
    assert cl.query(key).data != cl.query(key).data


def test_client_setup_errors():
    base = dict(num_items=32768, item_size_bits=1216 * 8, dim0=2048, cuckoo_seed=O.CUCKOO_SEED, num_hashes=2,
                key_size=32, value_size=1184, entry_size=1216, num_records=10_000)
    with pytest.raises(IA.InspireError, match="server keyword size must be non-zero"):
        IA.KeywordClient(dict(base, key_size=0))
    with pytest.raises(IA.InspireError, match="server value size must be non-zero"):
        IA.KeywordClient(dict(base, value_size=0))
    with pytest.raises(IA.InspireError, match="server entry size 100 is smaller than key_size 32 \\+ value_size 1184"):
        IA.KeywordClient(dict(base, entry_size=100))
    with pytest.raises(IA.InspireError, match="server bucket count 1024 is too small for the PIR scheme "
                                              "\\(needs at least 2048 records\\)"):
        IA.KeywordClient(dict(base, num_items=1024))
    cl = IA.KeywordClient(base)
    with pytest.raises(IA.InspireError, match="expected 32 keyword bytes, got 31"):
        cl.positions(bytes(31))
    # positions follow the handshake's cuckoo seed
    seed = bytes(range(16))
    key = bytes(range(32))
    assert IA.KeywordClient(dict(base, cuckoo_seed=seed)).positions(key) == O.cuckoo_positions(key, 32768, seed)


def test_framing_round_trip_and_limits():
    a, b = socket.socketpair()
    a.settimeout(30)
    b.settimeout(30)
    try:
        payload = bytes(range(256)) * 1000   # larger than the socket buffer: sent from a thread
        sender = threading.Thread(target=lambda: (wire.send_msg(a, wire.MSG_KEYWORD_QUERY, payload),
                                                  wire.send_msg(a, wire.MSG_HANDSHAKE, b"")))
        sender.start()
        assert wire.recv_msg(b) == (wire.MSG_KEYWORD_QUERY, payload)
        sender.join()
        assert wire.recv_msg(b) == (wire.MSG_HANDSHAKE, b"")
        # [8-byte BE length][type][payload]
        wire.send_msg(a, 0x06, b"xyz")
        assert b.recv(12) == b"\0\0\0\0\0\0\0\x03\x06xyz"
        # a declared length over 64 MiB is refused before the payload is read
        a.sendall((wire.MAX_MSG_SIZE + 1).to_bytes(8, "big") + b"\x05")
        with pytest.raises(wire.WireError, match="message too large"):
            wire.recv_msg(b)
        a.close()
        with pytest.raises(EOFError):
            wire.recv_msg(b)
    finally:
        a.close()
        b.close()


def test_handshake_json_layout():
    pp = dict(num_items=32768, item_size_bits=9728, dim0=2048, cuckoo_seed=O.CUCKOO_SEED, num_hashes=2,
              key_size=32, value_size=1184, entry_size=1216, num_records=10000)
    js = wire.public_params_to_json(pp)
    # serde_json's field order for PublicParams / KeywordPirHandshake (wire.rs:19-36), compact
    assert js == (b'{"num_items":32768,"item_size_bits":9728,"dim0":2048,"keyword_pir":{"cuckoo_seed":'
                  b'[222,173,190,239,202,254,186,190,1,35,69,103,137,171,205,239],"num_hashes":2,'
                  b'"key_size":32,"value_size":1184,"entry_size":1216,"num_records":10000}}')
    assert wire.public_params_from_json(js) == pp
    with pytest.raises(ValueError, match="server is not in keyword PIR mode"):
        wire.public_params_from_json(json.dumps({"num_items": 1, "item_size_bits": 16, "dim0": 2048,
                                                 "keyword_pir": None}).encode())


def test_server_cli_memory_estimate(tmp_path, capsys):
    # server.rs --estimate-memory: the plan and memory lines, then exit (no GPU needed)
    from inspire_amd import server as srvmod
    path = str(tmp_path / "ds.bin")
    IA.generate_dataset(path, "ml-kem-768", 3000, seed=3)
    assert srvmod.dataset_header(path) == ("ml-kem-768", 3000, 1184)
    assert srvmod.main(["--dataset", path, "--estimate-memory"]) == 0
    out = capsys.readouterr().out
    plan = IA.plan_shape(3000, 32, 1184)
    assert "startup GPU memory check: require %.2f GiB (%d bytes)" % (plan["peak_bytes"] / 2**30, plan["peak_bytes"]) in out
    assert ("estimate details: buckets %d, dim0 %d, db_rows %d, db_cols %d, instances %d"
            % (plan["num_items"], plan["dim0"], plan["db_rows"], plan["db_cols"], plan["instances"])) in out
