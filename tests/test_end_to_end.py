"""End-to-end decryption through the oracle: real server state (hint, prep-pack,
InspiRING offline precompute), a real client (packing keys, encrypted query
rows, RGSW selection bodies), the respond path, and the client's decrypt.

The reference's own invariant for its packing is test_packing_inspir
(packing.rs:3428-3526: pack LWE encryptions, decrypt, compare with the
messages); the end-to-end form of it is tests/e2e_mlkem.rs:27-38 (extract
recovers the stored key).  Both are restated here at the level of the answer
path: every stored value of the selected row and sub-column decrypts to the
raw DB value (mod p = 65535, so a stored 0xFFFF reads back as 0).  These tests
pin the semantics of the packing (rotation directions, bold_t vs bold_t_bar),
the post-process, the fold and the modulus switch, which the per-stage closed
forms alone do not fix."""
import numpy as np
import pytest

import oracle as O

N = O.N
DB_ROWS, INST, D = 2048, 4, 4


@pytest.fixture(scope="module")
def real_state():
    rng = np.random.default_rng(2024)
    raw = rng.integers(0, 65536, size=INST * N * DB_ROWS, dtype=np.uint16)
    raw.reshape(INST * N, DB_ROWS)[:16, 0] = 0xFFFF   # stored 0xFFFF words read back as 0
    enc = O.encode(raw, DB_ROWS, INST * N, D)
    bt, btb, bth, ahat = O.setup_state(enc, DB_ROWS, INST)
    return raw.reshape(INST * N, DB_ROWS), enc, (bt, btb, bth, ahat)


def expected(rawm, row, sub, inst=INST, d=D):
    c = inst // d
    return np.stack([rawm[(w * d + sub) * N:(w * d + sub + 1) * N, row] for w in range(c)]).astype(np.uint64) % 65535


@pytest.mark.parametrize("hamming", [0, O.HAMMING_WEIGHT])
def test_oracle_respond_decrypts_to_stored_values(real_state, hamming):
    rawm, enc, st = real_state
    sk = O.client_secret(11, hamming)
    targets = [(0, 0), (1999, 3)]
    y, z, rows, bodies = O.client_query(sk, DB_ROWS, D, targets, seed=5)
    blobs = O.respond_raw(enc, DB_ROWS, INST, D, *st, y, z, rows, bodies)
    for (row, sub), blob in zip(targets, blobs):
        got = O.decrypt_response(blob, sk)
        assert np.array_equal(got, expected(rawm, row, sub)), (row, sub)
    assert (expected(rawm, 0, 0)[0][:16] == 0).all()   # the planted 0xFFFF words


def test_oracle_packing_decrypts_gemv_messages(real_state):
    # Before the fold: packed_j = (a_hat_j, c_j + CRT(INTT(S_j))) decrypts to the
    # encoded column values of the selected row (packing.rs:3011-3020).
    rawm, enc, (bt, btb, bth, ahat) = real_state
    sk = O.client_secret(12, O.HAMMING_WEIGHT)
    y, z = O.client_packing_keys(sk, 6)
    row = 777
    qrow = O.client_query_row(DB_ROWS, row, sk, 6)
    cvec = O.gemv(qrow, enc, DB_ROWS, INST * N)
    P = O.post(O.packing(bt, btb, bth, y, z, INST))
    encm = enc.reshape(INST * N, DB_ROWS)
    for j in range(INST):
        ct = np.stack([ahat[j * N:(j + 1) * N], cvec[j * N:(j + 1) * N] + P[j]])
        got = O.client_decrypt_raw(ct, sk)
        assert np.array_equal(got, encm[j * N:(j + 1) * N, row].astype(np.uint64)), j


@pytest.mark.parametrize("mistake", ["swap_bold_t", "zero_bold_t_hat", "shift_a_hat", "wrong_z_body"])
def test_decrypt_detects_planted_state_mistakes(real_state, mistake):
    # Each plausible mistake in the packing's inputs breaks decryption, so the
    # tests above would catch it.
    rawm, enc, (bt, btb, bth, ahat) = real_state
    sk = O.client_secret(13, O.HAMMING_WEIGHT)
    y, z, rows, bodies = O.client_query(sk, DB_ROWS, D, [(42, 2)], seed=7)
    if mistake == "swap_bold_t":
        bt, btb = btb, bt
    elif mistake == "zero_bold_t_hat":
        bth = np.zeros_like(bth)
    elif mistake == "shift_a_hat":
        ahat = np.roll(ahat, 1)
    else:
        z = O.client_packing_keys(sk, 8)[0]   # y in place of z
    blobs = O.respond_raw(enc, DB_ROWS, INST, D, bt, btb, bth, ahat, y, z, rows, bodies)
    got = O.decrypt_response(blobs[0], sk)
    assert (got != expected(rawm, 42, 2)).sum() > N // 2


def test_packing_key_bodies_are_key_switching_keys():
    # generate_ksk_body (packing.rs:3272-3294): body_k + s*mask_k - tau(s)*2^(19k)
    # is the small noise e_k, for both y (tau = X -> X^5, mask W) and z (X -> X^-1, mask V).
    sk = O.client_secret(14, O.HAMMING_WEIGHT)
    assert (sk == 1).sum() == O.HAMMING_WEIGHT and (sk == O.Q - 1).sum() == O.HAMMING_WEIGHT
    y, z = O.client_packing_keys(sk, 9)
    w_seed, v_seed = [7] * 32, [8, 8] + [7] * 30
    for body, seed, t in ((y, w_seed, 5), (z, v_seed, 2 * N - 1)):
        mask = O.chacha20_u64(seed, 3 * 2 * N).reshape(3, 2, N)
        mask[:, 0] %= O.Q0
        mask[:, 1] %= O.Q1
        tau = O.automorph_raw(sk, t, O.Q)
        for k in range(3):
            res = []
            for c, q in enumerate((O.Q0, O.Q1)):
                b = (body[k] >> np.uint64(32 * c)) & np.uint64(0xFFFFFFFF)
                s_ntt = O.ntt_forward(np.stack([sk % O.Q0, sk % O.Q1])[None])[0][c]
                t_ntt = O.ntt_forward(np.stack([tau % O.Q0, tau % O.Q1])[None])[0][c]
                v = [(int(b[i]) + int(s_ntt[i]) * int(mask[k, c, i]) - int(t_ntt[i]) * (1 << (19 * k))) % q
                     for i in range(N)]
                res.append(np.array(v, dtype=np.uint64))
            e = O.ntt_inverse(np.stack(res)[None])[0]
            e0 = np.where(e[0] > O.Q0 // 2, e[0].astype(np.int64) - O.Q0, e[0].astype(np.int64))
            e1 = np.where(e[1] > O.Q1 // 2, e[1].astype(np.int64) - O.Q1, e[1].astype(np.int64))
            assert np.array_equal(e0, e1) and np.abs(e0).max() <= 65, (t, k)
