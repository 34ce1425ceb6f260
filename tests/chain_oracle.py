"""Oracle answers for single fold chains of a server filled by
fill_synthetic(seed).

The counter-based generator gives any slice of the synthetic state directly
(the DB at index col * db_rows + row, bold_t / bold_t_bar at
output * 3069 * 2048 + ..., bold_t_hat and a_hat likewise), so the oracle can
recompute chain w of a response -- its d packed outputs, the GEMV over its
d * 2048 columns, the fold and the modulus switch (keyword.rs:543-581,
modulus_switch.rs:45-73) -- at any database size without the rest of the
state.  Test helper only.
"""
import numpy as np

import oracle as O

N = O.N


def chain_blobs(seed, db_rows, d, w, y, z, rows, bodies):
    """Chain w's 12,288 response bytes for each hash query of one keyword query."""
    j0 = w * d
    bt = O.synth(seed, O.STREAM_BT, d * O.NINNER * N, O.KIND_CONDENSED, start=j0 * O.NINNER * N)
    btb = O.synth(seed, O.STREAM_BTB, d * O.NINNER * N, O.KIND_CONDENSED, start=j0 * O.NINNER * N)
    bth = O.synth(seed, O.STREAM_BTH, d * 3 * N, O.KIND_CONDENSED, start=j0 * 3 * N)
    ahat = O.synth(seed, O.STREAM_AHAT, d * N, O.KIND_MODQ, start=j0 * N).reshape(d, N)
    P = O.post(O.packing(bt, btb, bth, y, z, d))
    del bt, btb
    c = O.gemv_synth(rows, seed, O.STREAM_DB, db_rows, j0 * N, d * N)
    neg_a, _ = O.rgsw_neg_a_ntt()
    out = []
    for h in range(rows.shape[0]):
        ct = O.fold_chain(ahat, c[h].reshape(d, N) + P, neg_a, bodies[h])
        out.append(O.switch_and_keep(ct))
    return out


def split_response(resp):
    """Per-hash-query blobs of a serialized response (wire.rs:708-726)."""
    n = int.from_bytes(resp[:4], "little")
    out, o = [], 4
    for _ in range(n):
        ln = int.from_bytes(resp[o:o + 4], "little")
        out.append(resp[o + 4:o + 4 + ln])
        o += 4 + ln
    return out


def chain_of(blob, w):
    return blob[w * O.RESP_BYTES:(w + 1) * O.RESP_BYTES]


def stage_queries(torch, serialized):
    """Serialized queries back to back in one device buffer at a 256-byte
    stride: (device buffer, query length, stride)."""
    qlen = len(serialized[0])
    stride = (qlen + 16 + 255) // 256 * 256
    host = np.zeros(stride * len(serialized), dtype=np.uint8)
    for i, b in enumerate(serialized):
        host[i * stride:i * stride + qlen] = np.frombuffer(b, dtype=np.uint8)
    return torch.from_numpy(host).cuda(), qlen, stride
