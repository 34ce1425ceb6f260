"""setup_keyword_server end to end on the GPU, the reference's
tests/e2e_mlkem.rs:8-38 at its shape: a PQKEYV2 file of 10,000 ML-KEM-768
public keys from DatasetGenerator seed 7 (ids = SHA3-256 as dataset.rs:36-46),
the server built from it (cuckoo table, raw DB, encode and offline
precomputation on the device), a client query for present ids, and extract
returning the stored public key, whose id matches (public_key_matches_id).
ML-KEM-512 and ML-KEM-1024 datasets (832- and 1600-byte entries) likewise.
Two cases of the reference's own behaviour are checked as such: an absent id
gives None, and an item whose 608 columns run past the end of a
2048-coefficient polynomial cannot be extracted (the reference's slice runs
past the selected polynomial's coefficients)."""
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def keyword_server(tmp_path_factory):
    import torch
    assert torch.cuda.is_available()
    say("setup_keyword_server: DatasetGenerator(ml-kem-768, 10,000, seed 7)")
    path = tmp_path_factory.mktemp("ds") / "keys.pqkeyv2"
    IA.generate_dataset(str(path), "ml-kem-768", 10_000, seed=7)
    recs = IA.generate_records("ml-kem-768", 10_000, seed=7)
    srv, pp = IA.setup_keyword_server(str(path))
    plan = IA.plan_shape(10_000, 32, 1184)
    yield srv, pp, plan, recs
    srv.close()


def test_public_params(keyword_server):
    srv, pp, plan, recs = keyword_server
    assert pp["num_items"] == 32768 and pp["dim0"] == 2048 and pp["num_hashes"] == 2
    assert pp["key_size"] == 32 and pp["value_size"] == 1184 and pp["entry_size"] == 1216
    assert pp["item_size_bits"] == 1216 * 8 and pp["num_records"] == 10_000
    assert pp["cuckoo_seed"] == O.CUCKOO_SEED
    assert (plan["db_rows"], plan["instances"], plan["interpolate_degree"]) == (2048, 8, 8)


def _usable(client, rid):
    # neither straddling nor holding a 0xFFFF word (read back as 0, p = 65535)
    pos = client.positions(rid)
    return all(not client.straddles(b) for b in pos)


@pytest.mark.parametrize("hamming", [0, O.HAMMING_WEIGHT])
def test_extract_returns_stored_public_key(keyword_server, hamming):
    srv, pp, plan, recs = keyword_server
    client = O.KeywordClient(pp, plan, O.client_secret(99 + hamming, hamming))
    rng = np.random.default_rng(2026 + hamming)
    done = 0
    for idx in rng.permutation(len(recs)):
        rid, pk = recs[idx]
        words = np.frombuffer(rid + pk, dtype=">u2")
        if not _usable(client, rid) or (words == 0xFFFF).any():
            continue
        say("extract record", int(idx))
        pos, q = client.query(rid, seed=int(idx))
        resp = srv.respond_serialized(q)
        got = client.extract(rid, pos, resp)
        assert got == pk, int(idx)
        assert O.public_key_id(got) == rid
        done += 1
        if done == 4:
            break
    assert done == 4


@pytest.mark.parametrize("kem,count", [("ml-kem-512", 6000), ("ml-kem-1024", 6000)])
def test_other_kems_extract(tmp_path, kem, count):
    say("setup_keyword_server: DatasetGenerator(%s, %d, seed 17)" % (kem, count))
    path = tmp_path / "keys.pqkeyv2"
    IA.generate_dataset(str(path), kem, count, seed=17)
    recs = IA.generate_records(kem, count, seed=17)
    srv, pp = IA.setup_keyword_server(str(path))
    pk_len = IA.KEM_PUBLIC_KEY_LEN[kem]
    assert pp["value_size"] == pk_len and pp["num_records"] == count
    plan = IA.plan_shape(count, 32, pk_len)
    client = O.KeywordClient(pp, plan, O.client_secret(7, O.HAMMING_WEIGHT))
    done = 0
    for idx in np.random.default_rng(count).permutation(count):
        rid, pk = recs[idx]
        if not _usable(client, rid) or (np.frombuffer(rid + pk, dtype=">u2") == 0xFFFF).any():
            continue
        pos, q = client.query(rid, seed=int(idx))
        got = client.extract(rid, pos, srv.respond_serialized(q))
        assert got == pk, int(idx)
        assert O.public_key_id(got, kem) == rid
        done += 1
        if done == 2:
            break
    assert done == 2
    srv.close()


def test_absent_and_straddling_ids(keyword_server):
    srv, pp, plan, recs = keyword_server
    client = O.KeywordClient(pp, plan, O.client_secret(5, O.HAMMING_WEIGHT))
    absent = bytes(range(32))
    pos, q = client.query(absent, seed=1)
    assert client.extract(absent, pos, srv.respond_serialized(q)) is None
    # a present id whose item straddles a polynomial boundary in both of its
    # buckets is not recoverable through extract
    for idx, (rid, pk) in enumerate(recs):
        if all(client.straddles(b) for b in client.positions(rid)):
            pos, q = client.query(rid, seed=idx)
            assert client.extract(rid, pos, srv.respond_serialized(q)) is None
            break


def test_million_key_dataset_extracts(tmp_path):
    # The benchmark's shape with real state: a PQKEYV2 file of 1,000,000
    # 1184-byte keys, setup_keyword_server (cuckoo table, raw DB, device
    # encode and offline precomputation at 352 outputs), then queries for
    # present ids through the full 11-chain response and extract.
    import hashlib
    rng = np.random.default_rng(11)
    n, kl = 1_000_000, 1184
    say("1M-key dataset: generating")
    keys = rng.integers(0, 256, size=(n, kl), dtype=np.uint8)
    dom = b"pq-key:v1:ml-kem-768"
    ids = np.frombuffer(b"".join(hashlib.sha3_256(dom + keys[i].tobytes()).digest() for i in range(n)),
                        dtype=np.uint8).reshape(n, 32)
    path = tmp_path / "million.pqkeyv2"
    with open(path, "wb") as f:
        f.write(b"PQKEYV2\0" + bytes([10]) + b"ml-kem-768" + n.to_bytes(4, "little") + kl.to_bytes(2, "little"))
        f.write(np.concatenate([ids, keys], axis=1).tobytes())
    say("1M-key dataset: setup_keyword_server")
    srv, pp = IA.setup_keyword_server(str(path))
    plan = IA.plan_shape(n, 32, kl)
    assert pp["num_items"] == 2_097_152 and pp["num_records"] == n
    assert (plan["instances"], plan["interpolate_degree"], plan["c"]) == (352, 32, 11)
    srv.set_packing_kernel(IA.PACKING_MFMA)
    client = O.KeywordClient(pp, plan, O.client_secret(1234, O.HAMMING_WEIGHT))
    done = 0
    for idx in rng.permutation(n)[:200]:
        rid, pk = ids[idx].tobytes(), keys[idx].tobytes()
        if any(client.straddles(b) for b in client.positions(rid)) or (
                np.frombuffer(rid + pk, dtype=">u2") == 0xFFFF).any():
            continue
        say("1M-key dataset: extract record", int(idx))
        pos, q = client.query(rid, seed=int(idx))
        assert client.extract(rid, pos, srv.respond_serialized(q)) == pk, int(idx)
        done += 1
        if done == 3:
            break
    assert done == 3
    srv.close()


def _usable_records(keyword_server, n, seed):
    srv, pp, plan, recs = keyword_server
    checker = O.KeywordClient(pp, plan, O.client_secret(1, 0))
    rng = np.random.default_rng(seed)
    out = []
    for idx in rng.permutation(len(recs)):
        rid, pk = recs[idx]
        words = np.frombuffer(rid + pk, dtype=">u2")
        if _usable(checker, rid) and not (words == 0xFFFF).any():
            out.append((int(idx), rid, pk))
            if len(out) == n:
                break
    return out


@pytest.mark.parametrize("hamming", [0, O.HAMMING_WEIGHT])
def test_library_client_extracts_stored_public_key(keyword_server, hamming):
    # the library's KeywordClient (inspire_client_*) against the GPU server
    srv, pp, plan, recs = keyword_server
    client = IA.KeywordClient(pp, hamming=hamming)
    for idx, rid, pk in _usable_records(keyword_server, 4, 77 + hamming):
        say("library client extract record", idx, "hamming", hamming)
        q = client.query(rid)
        assert q.positions == O.cuckoo_positions(rid, pp["num_items"])
        got = client.extract_from_serialized_response(rid, q, srv.respond_serialized(q.data))
        assert got == pk and O.public_key_id(got) == rid, idx
    # an id stored nowhere (not the all-zero id: empty buckets hold zero
    # entries, so that one matches them, in the reference as here)
    absent = bytes(range(1, 33))
    q = client.query(absent)
    assert client.extract(absent, q, srv.respond_serialized(q.data)) is None


def test_tcp_server_answers_concurrent_clients(keyword_server):
    # src/bin/server.rs + client.rs over 127.0.0.1: the handshake, framed queries
    # from four connections at once, answered through one batching engine
    import threading
    from inspire_amd import client as cli
    from inspire_amd import server as srvmod
    srv, pp, plan, recs = keyword_server
    picks = _usable_records(keyword_server, 8, 5150)
    say("tcp server with", len(picks), "ids over 4 connections")
    tcp = srvmod.KeywordTcpServer(srv, pp, listen="127.0.0.1:0", max_batch=8)
    addr = tcp.start()
    results, errors = {}, []

    def worker(part):
        try:
            for idx, rid, pk in part:
                results[idx] = (cli.fetch(addr, rid, hamming=O.HAMMING_WEIGHT)[0], pk)
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    try:
        threads = [threading.Thread(target=worker, args=(picks[i::4],)) for i in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not errors, errors
        assert len(results) == len(picks) and all(got == pk for got, pk in results.values())
        assert cli.fetch(addr, bytes(range(1, 33)))[0] is None
        assert tcp.answered == len(picks) + 1
        assert tcp.engine.stats()["queries"] == len(picks) + 1
    finally:
        tcp.stop()


def test_server_and_client_binaries(tmp_path):
    # python -m inspire_amd.server / inspire_amd.client as processes, as the
    # reference's bin/server.rs and bin/client.rs are run (tests/e2e_tcp_mlkem.rs)
    import os
    import re
    import subprocess
    import time
    path = str(tmp_path / "keys.pqkeyv2")
    IA.generate_dataset(path, "ml-kem-768", 3000, seed=11)
    recs = IA.generate_records("ml-kem-768", 3000, seed=11)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([IA.__path__[0] + "/..", env.get("PYTHONPATH", "")])
    say("server process on 127.0.0.1:0")
    proc = subprocess.Popen([sys.executable, "-u", "-m", "inspire_amd.server", "--dataset", path,
                             "--listen", "127.0.0.1:0"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            text=True, env=env)
    try:
        addr, lines, t0 = None, [], time.time()
        while addr is None and time.time() - t0 < 240:
            line = proc.stdout.readline()
            if not line:
                break
            lines.append(line)
            m = re.search(r"serving 3000 ml-kem-768 public keys on (\S+)", line)
            if m:
                addr = m.group(1)
        assert addr, "".join(lines)
        assert any("startup GPU memory check: require" in x for x in lines)
        checker = O.KeywordClient({"num_items": 8192, "item_size_bits": 9728, "dim0": 2048,
                                   "cuckoo_seed": O.CUCKOO_SEED, "key_size": 32, "value_size": 1184,
                                   "entry_size": 1216}, IA.plan_shape(3000, 32, 1184), O.client_secret(1, 0))
        done = 0
        for idx in range(len(recs)):
            rid, pk = recs[idx]
            if not _usable(checker, rid) or (np.frombuffer(rid + pk, dtype=">u2") == 0xFFFF).any():
                continue
            say("client process for record", idx)
            out = subprocess.run([sys.executable, "-m", "inspire_amd.client", "--server", addr, "--id", rid.hex()],
                                 capture_output=True, text=True, env=env, timeout=120)
            assert out.returncode == 0, out.stderr
            assert ("public key: %s" % pk.hex()) in out.stdout and ("query id: %s" % rid.hex()) in out.stdout
            done += 1
            if done == 2:
                break
        out = subprocess.run([sys.executable, "-m", "inspire_amd.client", "--server", addr, "--id",
                              bytes(range(1, 33)).hex()], capture_output=True, text=True, env=env, timeout=120)
        assert out.returncode == 0 and "not found" in out.stdout, out.stdout + out.stderr
        # the all-zero id matches an empty bucket's zero entry, as in the reference
        out = subprocess.run([sys.executable, "-m", "inspire_amd.client", "--server", addr, "--id", "00" * 32],
                             capture_output=True, text=True, env=env, timeout=120)
        assert out.returncode == 0 and ("public key: " + "00" * 1184) in out.stdout, out.stdout + out.stderr
    finally:
        proc.terminate()
        try:
            proc.wait(timeout=30)
        except subprocess.TimeoutExpired:
            proc.kill()
            proc.wait()
