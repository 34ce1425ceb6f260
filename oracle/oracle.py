"""ctypes wrapper around the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It is the checker, never the thing measured or
shipped.  Each wrapper names the reference function the C code restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")          # x86-64-v3 (AVX2)
LIB_PATH_V4 = os.path.join(HERE, "liboracle_v4.so")    # x86-64-v4 (AVX-512)
_V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")

N = 2048
TEXP = 3
NROT = N // 2 - 1
NINNER = NROT * TEXP
Q0 = 268369921
Q1 = 249561089
Q = Q0 * Q1
RESP_BYTES = 12288

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "all"], check=True)


def host_info():
    """CPU model, hardware threads, RAM and the ISA level of the loaded oracle."""
    model, flags = "", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and not model:
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    ram = 0
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                ram = int(line.split()[1]) * 1024
    except OSError:
        pass
    return {"cpu_model": model, "hw_threads": os.cpu_count(), "ram_bytes": ram,
            "avx512": all(f in flags for f in _V4_FLAGS), "oracle_isa": lib_isa()}


def _has_v4():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return all(f in line.split() for f in _V4_FLAGS)
    except OSError:
        pass
    return False


_isa = None


def lib_isa():
    lib()
    return _isa


def lib():
    """The oracle library: the AVX-512 build on hosts that have AVX-512, else AVX2."""
    global _lib, _isa
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        path, _isa = LIB_PATH, "x86-64-v3 (AVX2)"
        if _has_v4() and os.path.exists(LIB_PATH_V4):
            path, _isa = LIB_PATH_V4, "x86-64-v4 (AVX-512)"
        _lib = ctypes.CDLL(path)
        _lib.orc_synth_word.restype = ctypes.c_uint64
        _lib.orc_synth_word.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        _lib.orc_read_bits.restype = ctypes.c_uint64
        _lib.orc_serialize_query.restype = ctypes.c_size_t
        _lib.orc_deserialize_query.restype = ctypes.c_int
        _lib.orc_serialize_response.restype = ctypes.c_size_t
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def constants():
    """core/params.rs:258-270 + arith.rs:117-131: Barrett and CRT constants."""
    out = np.zeros(16, dtype=np.uint64)
    lib().orc_constants(_p(out))
    keys = ["cr0_q0", "cr1_q0", "cr0_q1", "cr1_q1", "cr0_q", "cr1_q", "mod0_inv_mod1",
            "mod1_inv_mod0", "modulus", "root0", "root1"]
    return {k: int(out[i]) for i, k in enumerate(keys)}


def ntt_tables():
    """ntt.rs:74-108 build_ntt_tables -> [2][4][N]."""
    out = np.zeros((2, 4, N), dtype=np.uint64)
    lib().orc_ntt_tables(_p(out))
    return out


def ntt_forward(polys):
    """ntt.rs:211-357 (restated in scalar form). polys: [..., 2, N] canonical."""
    a = np.ascontiguousarray(polys, dtype=np.uint64).copy()
    cnt = a.size // (2 * N)
    lib().orc_ntt_forward(_p(a), ctypes.c_size_t(cnt))
    return a


def ntt_inverse(polys):
    """ntt.rs:575-728."""
    a = np.ascontiguousarray(polys, dtype=np.uint64).copy()
    cnt = a.size // (2 * N)
    lib().orc_ntt_inverse(_p(a), ctypes.c_size_t(cnt))
    return a


def crt_compose(x, y):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    y = np.ascontiguousarray(y, dtype=np.uint64)
    out = np.zeros_like(x)
    lib().orc_crt_compose(_p(x), _p(y), _p(out), ctypes.c_size_t(x.size))
    return out


def barrett_u128(vals, cr0, cr1, m):
    lo = np.array([v & ((1 << 64) - 1) for v in vals], dtype=np.uint64)
    hi = np.array([v >> 64 for v in vals], dtype=np.uint64)
    out = np.zeros(len(vals), dtype=np.uint64)
    lib().orc_barrett_u128(_p(lo), _p(hi), _p(out), ctypes.c_size_t(len(vals)),
                           ctypes.c_uint64(cr0), ctypes.c_uint64(cr1), ctypes.c_uint64(m))
    return out


def barrett_raw_u64(vals, cr1, m):
    x = np.ascontiguousarray(vals, dtype=np.uint64)
    out = np.zeros_like(x)
    lib().orc_barrett_raw_u64(_p(x), _p(out), ctypes.c_size_t(x.size), ctypes.c_uint64(cr1),
                              ctypes.c_uint64(m))
    return out


def automorph_table(t):
    """packing.rs:1423-1468 (closed form)."""
    out = np.zeros(N, dtype=np.uint32)
    lib().orc_automorph_table(ctypes.c_uint32(t), _p(out))
    return out


def automorph_raw(a, t, modulus):
    """poly.rs automorph_poly."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros_like(a)
    lib().orc_automorph_raw(_p(a), _p(out), ctypes.c_uint32(t), ctypes.c_uint64(modulus))
    return out


def gadget_invert(inp, mx, modulus_log2=56):
    """gadget.rs:35-65. inp [rdim][N] -> [mx][N]."""
    inp = np.ascontiguousarray(inp, dtype=np.uint64)
    rdim = inp.shape[0]
    out = np.zeros((mx, N), dtype=np.uint64)
    lib().orc_gadget_invert(_p(inp), _p(out), ctypes.c_int(rdim), ctypes.c_int(mx),
                            ctypes.c_uint64(modulus_log2))
    return out


def rescale(vals, inp_mod, out_mod):
    """arith.rs:436-451."""
    a = np.ascontiguousarray(vals, dtype=np.uint64)
    out = np.zeros_like(a)
    lib().orc_rescale(_p(a), _p(out), ctypes.c_size_t(a.size), ctypes.c_uint64(inp_mod),
                      ctypes.c_uint64(out_mod))
    return out


def chacha20_block(key, words12_15):
    k = np.frombuffer(bytes(key), dtype=np.uint8).copy()
    w = np.array(words12_15, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().orc_chacha20_block(_p(k), _p(w), _p(out))
    return out


def chacha20_u64(seed, n):
    s = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_chacha20_u64(_p(s), _p(out), ctypes.c_size_t(n))
    return out


def rgsw_neg_a_ntt():
    """keyword.rs:546-555: NTT(-a) [2][N] and a (raw mod q)."""
    out = np.zeros((2, N), dtype=np.uint64)
    a = np.zeros(N, dtype=np.uint64)
    lib().orc_rgsw_neg_a_ntt(_p(out), _p(a))
    return out, a


def gemv(a, db, db_rows, db_cols):
    """gemv_kernel.cu:46-103 / kernel.rs:139-237."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    db = np.ascontiguousarray(db, dtype=np.uint16)
    c = np.zeros(db_cols, dtype=np.uint64)
    lib().orc_gemv(_p(c), _p(a), _p(db), ctypes.c_int64(db_rows), ctypes.c_int64(db_cols))
    return c


def gemv_synth(rows, seed, stream, db_rows, j0, ncols):
    """gemv for each query row of `rows` [nq][db_rows] over columns
    [j0, j0 + ncols) of the synthetic DB (synth kind KIND_U16 of stream
    `stream`), without a host copy of the DB -> [nq][ncols]."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    nq = rows.shape[0]
    assert rows.shape[1] == db_rows and 1 <= nq <= 8 and db_rows <= (1 << 19)
    c = np.zeros((nq, ncols), dtype=np.uint64)
    lib().orc_gemv_synth(_p(c), _p(rows), ctypes.c_int(nq), ctypes.c_uint64(seed), ctypes.c_uint32(stream),
                         ctypes.c_int64(db_rows), ctypes.c_int64(j0), ctypes.c_int64(ncols))
    return c


def packing(bold_t, bold_t_bar, bold_t_hat, y_body, z_body, instances, j0=0, j1=None):
    """packing_online_kernel.cu:18-92 -> S [j1-j0][2][N]."""
    if j1 is None:
        j1 = instances
    S = np.zeros((j1 - j0, 2, N), dtype=np.uint64)
    lib().orc_packing(_p(S), _p(bold_t), _p(bold_t_bar), _p(bold_t_hat), _p(y_body), _p(z_body),
                      ctypes.c_int64(instances), ctypes.c_int64(j0), ctypes.c_int64(j1))
    return S


def post(S):
    """packing_online.rs:359-398: CRT(INTT(S)) -> [count][N] mod q."""
    S = np.ascontiguousarray(S, dtype=np.uint64)
    cnt = S.shape[0]
    P = np.zeros((cnt, N), dtype=np.uint64)
    lib().orc_post(_p(P), _p(S), ctypes.c_int64(cnt))
    return P


def external_product(sum2, neg_a, body):
    sum2 = np.ascontiguousarray(sum2, dtype=np.uint64)
    out = np.zeros((2, N), dtype=np.uint64)
    lib().orc_external_product(_p(out), _p(sum2), _p(np.ascontiguousarray(neg_a, dtype=np.uint64)),
                               _p(np.ascontiguousarray(body, dtype=np.uint64)))
    return out


def fold_chain(packed_a, packed_b, neg_a, body):
    """keyword.rs:558-573 for one chain; packed_a/b: [d][N]."""
    packed_a = np.ascontiguousarray(packed_a, dtype=np.uint64)
    packed_b = np.ascontiguousarray(packed_b, dtype=np.uint64)
    out = np.zeros((2, N), dtype=np.uint64)
    lib().orc_fold_chain(_p(out), _p(packed_a), _p(packed_b), ctypes.c_int(packed_a.shape[0]),
                         _p(np.ascontiguousarray(neg_a, dtype=np.uint64)),
                         _p(np.ascontiguousarray(body, dtype=np.uint64)))
    return out


def switch_and_keep(ct):
    """modulus_switch.rs:45-73 -> 12288 bytes."""
    ct = np.ascontiguousarray(ct, dtype=np.uint64)
    out = np.zeros(RESP_BYTES, dtype=np.uint8)
    lib().orc_switch_and_keep(_p(ct), _p(out))
    return out.tobytes()


def respond_raw(db, db_rows, instances, d, bold_t, bold_t_bar, bold_t_hat, a_hat,
                y_body, z_body, rows, bodies):
    """keyword.rs:534-628 up to the per-hash response blobs (list of bytes)."""
    nq = rows.shape[0]
    c = instances // d
    out = np.zeros(nq * c * RESP_BYTES, dtype=np.uint8)
    rc = lib().orc_respond_raw(_p(out), _p(db), ctypes.c_int64(db_rows), ctypes.c_int64(instances),
                               ctypes.c_int(d), _p(bold_t), _p(bold_t_bar), _p(bold_t_hat),
                               _p(a_hat), ctypes.c_int(nq), _p(y_body), _p(z_body),
                               _p(np.ascontiguousarray(rows)), _p(np.ascontiguousarray(bodies)))
    if rc != 0:
        raise ValueError("orc_respond_raw failed: %d" % rc)
    blob = out.tobytes()
    per = c * RESP_BYTES
    return [blob[i * per:(i + 1) * per] for i in range(nq)]


def serialize_query(y_body, z_body, rows, bodies):
    """wire.rs:542-615."""
    nq, db_rows = rows.shape
    cap = 1 + ((TEXP * N * 2 + nq * (db_rows + 6 * N)) * 56 + 63) // 64 * 8 + 16
    out = np.zeros(cap, dtype=np.uint8)
    n = lib().orc_serialize_query(_p(out), ctypes.c_size_t(cap), ctypes.c_int(nq),
                                  ctypes.c_int64(db_rows), _p(y_body), _p(z_body),
                                  _p(np.ascontiguousarray(rows)), _p(np.ascontiguousarray(bodies)))
    return out[:n].tobytes()


def deserialize_query(buf, db_rows, max_q=8):
    """wire.rs:618-705."""
    b = np.frombuffer(buf, dtype=np.uint8).copy()
    y = np.zeros((TEXP, N), dtype=np.uint64)
    z = np.zeros((TEXP, N), dtype=np.uint64)
    rows = np.zeros((max_q, db_rows), dtype=np.uint64)
    bodies = np.zeros((max_q, 6, 2, N), dtype=np.uint64)
    nq = lib().orc_deserialize_query(_p(b), ctypes.c_size_t(len(buf)), ctypes.c_int64(db_rows),
                                     ctypes.c_int(max_q), _p(y), _p(z), _p(rows), _p(bodies))
    if nq < 0:
        raise ValueError("malformed keyword query")
    return y, z, rows[:nq], bodies[:nq]


def serialize_response(responses, stash):
    """wire.rs:708-726."""
    out = bytearray()
    out += len(responses).to_bytes(4, "little")
    for r in responses:
        out += len(r).to_bytes(4, "little") + r
    out += len(stash).to_bytes(4, "little")
    for e in stash:
        out += len(e).to_bytes(4, "little") + e
    return bytes(out)


# ---------------------------------------------------------------------------
# Synthetic inputs (same counter-based generator as the HIP library)
# ---------------------------------------------------------------------------

KIND_CONDENSED, KIND_MODQ, KIND_U16, KIND_CRT, KIND_SMALL = 0, 1, 2, 3, 4
STREAM_DB, STREAM_BT, STREAM_BTB, STREAM_BTH, STREAM_AHAT = 1, 2, 3, 4, 5
STREAM_Y, STREAM_Z, STREAM_ROW, STREAM_BODY = 16, 17, 18, 19


def synth(seed, stream, count, kind, start=0):
    dt = np.uint16 if kind == KIND_U16 else np.uint64
    out = np.zeros(count, dtype=dt)
    lib().orc_synth_fill(_p(out), ctypes.c_uint64(seed), ctypes.c_uint32(stream),
                         ctypes.c_uint64(start), ctypes.c_uint64(count), ctypes.c_int(kind))
    return out


def synth_state(seed, db_rows, instances):
    """Synthetic server state: encoded DB, bold_t/bar/hat, a_hat."""
    db_cols = instances * N
    db = synth(seed, STREAM_DB, db_rows * db_cols, KIND_U16)
    bt = synth(seed, STREAM_BT, instances * NINNER * N, KIND_CONDENSED)
    btb = synth(seed, STREAM_BTB, instances * NINNER * N, KIND_CONDENSED)
    bth = synth(seed, STREAM_BTH, instances * TEXP * N, KIND_CONDENSED)
    ahat = synth(seed, STREAM_AHAT, instances * N, KIND_MODQ)
    return db, bt, btb, bth, ahat


def synth_query(seed, db_rows, nq=2, small=False):
    """Synthetic keyword query (keys + nq per-hash queries)."""
    y = synth(seed, STREAM_Y, TEXP * N, KIND_CONDENSED).reshape(TEXP, N)
    z = synth(seed, STREAM_Z, TEXP * N, KIND_CONDENSED).reshape(TEXP, N)
    rows = synth(seed, STREAM_ROW, nq * db_rows, KIND_SMALL if small else KIND_CONDENSED)
    rows = rows.reshape(nq, db_rows)
    bodies = synth(seed, STREAM_BODY, nq * 6 * 2 * N, KIND_CRT).reshape(nq, 6, 2, N)
    return y, z, rows, bodies


def num_threads():
    return lib().orc_num_threads()


def encode(raw, db_rows, db_cols, d, gamma=N, modulus=65535):
    """encode_kernel.cu:18-228 (gpu_encode_db): interpolation of each DB row's
    column groups; raw and the result are column-major u16 [db_cols][db_rows]."""
    raw = np.ascontiguousarray(raw, dtype=np.uint16).reshape(-1)
    assert raw.size == db_rows * db_cols and db_cols % (d * gamma) == 0
    num_inv = pow(d, -1, modulus)
    out = np.zeros_like(raw)
    lib().orc_encode(_p(raw), _p(out), ctypes.c_int64(db_rows), ctypes.c_int64(db_cols), ctypes.c_int(d),
                     ctypes.c_int(gamma), ctypes.c_uint32(modulus), ctypes.c_uint32(num_inv))
    return out


# ---------------------------------------------------------------------------
# Client and server setup (producers of real respond-path state)
# ---------------------------------------------------------------------------

HAMMING_WEIGHT = 256  # client/base.rs:14
P_MOD = 65535


def client_secret(seed, hamming=HAMMING_WEIGHT):
    """gen_ternary_mat (client/base.rs:126-139); hamming=0 is the keyword
    client's zero key (Client::init, client/base.rs:370-389)."""
    sk = np.zeros(N, dtype=np.uint64)
    lib().orc_client_secret(ctypes.c_uint64(seed), ctypes.c_int(hamming), _p(sk))
    return sk


def client_packing_keys(sk, seed):
    """PackingKeys::init_full (packing.rs:1895-1944) -> y_body, z_body [3][N] condensed."""
    y = np.zeros((TEXP, N), dtype=np.uint64)
    z = np.zeros((TEXP, N), dtype=np.uint64)
    lib().orc_client_packing_keys(_p(y), _p(z), _p(np.ascontiguousarray(sk, dtype=np.uint64)),
                                  ctypes.c_uint64(seed))
    return y, z


def client_query_row(db_rows, target_row, sk, seed, stream=0x70):
    """generate_query_b_values + pack_query (client/online.rs:226-357, 26-52)."""
    out = np.zeros(db_rows, dtype=np.uint64)
    lib().orc_client_query_row(_p(out), ctypes.c_int64(db_rows), ctypes.c_int64(target_row),
                               _p(np.ascontiguousarray(sk, dtype=np.uint64)), ctypes.c_uint64(seed),
                               ctypes.c_uint32(stream))
    return out


def client_rgsw_body(sk, exponent, seed, stream=0x80):
    """ct_gsw_body (client/keyword.rs:150-178) -> [6][2][N] NTT residues."""
    out = np.zeros((6, 2, N), dtype=np.uint64)
    lib().orc_client_rgsw_body(_p(out), _p(np.ascontiguousarray(sk, dtype=np.uint64)),
                               ctypes.c_uint32(exponent), ctypes.c_uint64(seed), ctypes.c_uint32(stream))
    return out


def rgsw_exponent(sub_col, d):
    """client/keyword.rs:156-158: (2n * target_sub_col / d) mod 2n."""
    return (2 * N * sub_col // d) % (2 * N)


def client_query(sk, db_rows, d, targets, seed):
    """KeywordClient::query for explicit (target_row, target_sub_col) pairs, one
    per hash query (client/keyword.rs:147-206) -> (y, z, rows, bodies)."""
    y, z = client_packing_keys(sk, seed)
    rows = np.zeros((len(targets), db_rows), dtype=np.uint64)
    bodies = np.zeros((len(targets), 6, 2, N), dtype=np.uint64)
    for i, (row, sub) in enumerate(targets):
        rows[i] = client_query_row(db_rows, row, sk, seed, 0x70 + i)
        bodies[i] = client_rgsw_body(sk, rgsw_exponent(sub, d), seed, 0x80 + i)
    return y, z, rows, bodies


def setup_atilde(db, db_rows, c0, ncols):
    """Hint + prep_pack (server/engine.rs:733-800, packing.rs:1032-1104) ->
    NTT residues [ncols][2][N] of the packing inputs of columns c0 .. c0+ncols-1."""
    out = np.zeros((ncols, 2, N), dtype=np.uint64)
    lib().orc_setup_atilde(_p(out), _p(np.ascontiguousarray(db, dtype=np.uint16)), ctypes.c_int64(db_rows),
                           ctypes.c_int64(c0), ctypes.c_int64(ncols))
    return out


def setup_precomp(at, instances):
    """full_packing_with_preprocessing_offline (packing.rs:2477-2750) for
    `instances` outputs -> bold_t, bold_t_bar, bold_t_hat (condensed), a_hat."""
    bt = np.zeros(instances * NINNER * N, dtype=np.uint64)
    btb = np.zeros(instances * NINNER * N, dtype=np.uint64)
    bth = np.zeros(instances * TEXP * N, dtype=np.uint64)
    ahat = np.zeros(instances * N, dtype=np.uint64)
    lib().orc_setup_precomp(_p(np.ascontiguousarray(at, dtype=np.uint64)), ctypes.c_int64(instances),
                            _p(bt), _p(btb), _p(bth), _p(ahat))
    return bt, btb, bth, ahat


def setup_state(db_enc, db_rows, instances):
    """KeywordServer::setup's offline precomputation (keyword.rs:427-458) over an
    encoded DB, one output polynomial at a time -> (bold_t, bold_t_bar, bold_t_hat, a_hat)."""
    bt = np.zeros(instances * NINNER * N, dtype=np.uint64)
    btb = np.zeros(instances * NINNER * N, dtype=np.uint64)
    bth = np.zeros(instances * TEXP * N, dtype=np.uint64)
    ahat = np.zeros(instances * N, dtype=np.uint64)
    for t in range(instances):
        at = setup_atilde(db_enc, db_rows, t * N, N)
        o = setup_precomp(at, 1)
        bt[t * NINNER * N:(t + 1) * NINNER * N] = o[0]
        btb[t * NINNER * N:(t + 1) * NINNER * N] = o[1]
        bth[t * TEXP * N:(t + 1) * TEXP * N] = o[2]
        ahat[t * N:(t + 1) * N] = o[3]
    return bt, btb, bth, ahat


def client_decrypt(blob, sk):
    """recover_how_many + decrypt_ct_reg_measured (modulus_switch.rs:102-134,
    client/online.rs:118-136) of one 12288-byte chain blob -> [N] values mod p."""
    b = np.frombuffer(bytes(blob), dtype=np.uint8).copy()
    assert b.size == RESP_BYTES
    out = np.zeros(N, dtype=np.uint64)
    lib().orc_client_decrypt(_p(b), _p(np.ascontiguousarray(sk, dtype=np.uint64)), _p(out))
    return out


def client_decrypt_raw(ct, sk):
    """Decrypt of an unswitched RLWE [2][N] (a, b) mod q, rescaled to p."""
    out = np.zeros(N, dtype=np.uint64)
    lib().orc_client_decrypt_raw(_p(np.ascontiguousarray(ct, dtype=np.uint64)),
                                 _p(np.ascontiguousarray(sk, dtype=np.uint64)), _p(out))
    return out


def decrypt_response(blob, sk):
    """All chains of one hash query's response -> [c][N] values mod p."""
    c = len(blob) // RESP_BYTES
    return np.stack([client_decrypt(blob[w * RESP_BYTES:(w + 1) * RESP_BYTES], sk) for w in range(c)])


# ---------------------------------------------------------------------------
# Keyword client (client/keyword.rs) and PQKEYV2 datasets (dataset/dataset.rs)
# ---------------------------------------------------------------------------

CUCKOO_SEED = bytes([0xDE, 0xAD, 0xBE, 0xEF, 0xCA, 0xFE, 0xBA, 0xBE,
                     0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF])  # cuckoo.rs:27-29


def siphash24(k0, k1, msg):
    lib().orc_siphash24.restype = ctypes.c_uint64
    mb = np.frombuffer(bytes(msg) or b"\0", dtype=np.uint8).copy()
    return lib().orc_siphash24(ctypes.c_uint64(k0), ctypes.c_uint64(k1), _p(mb), ctypes.c_size_t(len(msg)))


def cuckoo_positions(key, num_buckets, seed=CUCKOO_SEED):
    """CuckooHash::hash for hash indices 0, 1 (cuckoo.rs:98-132)."""
    out = []
    for h in range(2):
        k0 = int.from_bytes(seed[:8], "little") ^ h
        k1 = int.from_bytes(seed[8:], "little") ^ (h << 32)
        out.append(siphash24(k0, k1, key) % num_buckets)
    return out


def public_key_id(public_key, kem_name="ml-kem-768"):
    """derive_public_key_id (dataset.rs:36-46): SHA3-256(b"pq-key:v1:" || kem || pk)."""
    import hashlib
    return hashlib.sha3_256(b"pq-key:v1:" + kem_name.encode() + bytes(public_key)).digest()


def write_pqkeyv2(path, public_keys, kem_name="ml-kem-768"):
    """save_dataset (dataset.rs:48-93) -> list of (id, public_key)."""
    recs = [(public_key_id(pk, kem_name), bytes(pk)) for pk in public_keys]
    with open(path, "wb") as f:
        f.write(b"PQKEYV2\0" + bytes([len(kem_name)]) + kem_name.encode())
        f.write(len(recs).to_bytes(4, "little") + len(recs[0][1]).to_bytes(2, "little"))
        for i, pk in recs:
            f.write(i + pk)
    return recs


class KeywordClient:
    """KeywordClient (client/keyword.rs:45-311) from public params and the
    server plan; sk: the client's secret (zero for the reference's client)."""

    def __init__(self, pp, plan, sk):
        self.pp, self.sk = pp, sk
        self.db_rows, self.d = plan["db_rows"], plan["interpolate_degree"]
        self.per = pp["num_items"] // self.db_rows
        self.ise = pp["item_size_bits"] // 16

    def positions(self, key):
        return cuckoo_positions(key, self.pp["num_items"], self.pp["cuckoo_seed"])

    def target(self, bucket):
        """build_query_for_index (client/keyword.rs:147-153) -> (row, sub_col)."""
        row = bucket // self.per
        col = (bucket % self.per) * self.ise
        return row, (col % (self.d * N)) // N

    def query(self, key, seed):
        pos = self.positions(key)
        y, z, rows, bodies = client_query(self.sk, self.db_rows, self.d, [self.target(b) for b in pos], seed)
        return pos, serialize_query(y, z, rows, bodies)

    def extract(self, key, positions, response):
        """extract (client/keyword.rs:232-299): the value, or None.  An item
        whose columns run past the end of its 2048-coefficient polynomial reads
        past the decrypted coefficients of the selected polynomial; the
        reference's slice then panics (one chain) or reads the next chain; here
        the short or wrong slice fails the key comparison and gives None."""
        ks, vs, es = self.pp["key_size"], self.pp["value_size"], self.pp["entry_size"]
        n = int.from_bytes(response[:4], "little")
        o, blobs = 4, []
        for _ in range(n):
            ln = int.from_bytes(response[o:o + 4], "little")
            blobs.append(response[o + 4:o + 4 + ln])
            o += 4 + ln
        m = int.from_bytes(response[o:o + 4], "little")
        o += 4
        stash = []
        for _ in range(m):
            ln = int.from_bytes(response[o:o + 4], "little")
            stash.append(response[o + 4:o + 4 + ln])
            o += 4 + ln
        for i, blob in enumerate(blobs):
            dec = decrypt_response(blob, self.sk).reshape(-1)
            col = (positions[i] % self.per) * self.ise
            off = (col // N // self.d) * N + col % N
            vals = dec[off:off + self.ise]
            entry = b"".join(int(v & 0xFFFF).to_bytes(2, "big") for v in vals)
            if len(entry) >= es and entry[:ks] == bytes(key):
                return entry[ks:ks + vs]
        for e in stash:
            if len(e) >= es and e[:ks] == bytes(key):
                return e[ks:ks + vs]
        return None

    def straddles(self, bucket):
        col = (bucket % self.per) * self.ise
        return col % N + self.ise > N
