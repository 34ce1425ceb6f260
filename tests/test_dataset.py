"""DatasetGenerator and PQKEYV2 (src/dataset/generator.rs:61-193, kem.rs:27-53,
writer.rs:17-79, dataset.rs:8-46), host side.

Pins:
- SHA3-256 / SHA3-512 / SHAKE128 / SHAKE256 against Python's hashlib;
- ML-KEM encapsulation keys against a restatement of FIPS 203 K-PKE.KeyGen
  written here on hashlib's SHA3 and SHAKE, whose NTT is pinned by its
  definition (NTT(f)[2i] + NTT(f)[2i+1] X = f mod X^2 - zeta^(2 BitRev7(i)+1))
  and whose base-case products are pinned by the negacyclic schoolbook product;
- the record seed against the oracle's ChaCha20 block (RFC 8439 vectors) keyed
  by a PCG32 restatement of rand_core's seed_from_u64;
- the file layout of writer.rs:38-45 and the ids of dataset.rs:36-46.
Byte parity with the ml-kem 0.2.3 crate, rand_core's seed_from_u64 and
rand_chacha's word positions rests on those published algorithms (no crate or
FIPS 203 known-answer file is in the image): parity unpinned beyond them.
"""
import hashlib
import os
import random
import struct

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

Q = 3329


def bitrev7(i):
    return int("{:07b}".format(i)[::-1], 2)


ZETA = [pow(17, bitrev7(i), Q) for i in range(128)]
GAMMA = [pow(17, 2 * bitrev7(i) + 1, Q) for i in range(128)]


def ntt(f):
    """FIPS 203 Algorithm 9."""
    f = list(f)
    i = 1
    length = 128
    while length >= 2:
        for start in range(0, 256, 2 * length):
            z = ZETA[i]
            i += 1
            for j in range(start, start + length):
                t = z * f[j + length] % Q
                f[j + length] = (f[j] - t) % Q
                f[j] = (f[j] + t) % Q
        length //= 2
    return f


def base_mul(fh, gh):
    """FIPS 203 Algorithms 11-12."""
    h = [0] * 256
    for i in range(128):
        a0, a1, b0, b1 = fh[2 * i], fh[2 * i + 1], gh[2 * i], gh[2 * i + 1]
        h[2 * i] = (a0 * b0 + a1 * b1 * GAMMA[i]) % Q
        h[2 * i + 1] = (a0 * b1 + a1 * b0) % Q
    return h


def sample_ntt(rho, j, i):
    """FIPS 203 Algorithm 7 on SHAKE128(rho || j || i)."""
    stream = hashlib.shake_128(rho + bytes([j, i])).digest(3 * 1024)
    out, t = [], 0
    while len(out) < 256:
        c0, c1, c2 = stream[t], stream[t + 1], stream[t + 2]
        t += 3
        d1, d2 = c0 + 256 * (c1 % 16), c1 // 16 + 16 * c2
        if d1 < Q:
            out.append(d1)
        if d2 < Q and len(out) < 256:
            out.append(d2)
    return out


def cbd(sigma, nonce, eta):
    """FIPS 203 Algorithm 8 of PRF_eta(sigma, N) = SHAKE256(sigma || N)."""
    b = hashlib.shake_256(sigma + bytes([nonce])).digest(64 * eta)
    bits = [(b[k >> 3] >> (k & 7)) & 1 for k in range(8 * len(b))]
    return [(sum(bits[2 * i * eta:2 * i * eta + eta]) - sum(bits[2 * i * eta + eta:2 * i * eta + 2 * eta])) % Q
            for i in range(256)]


def kpke_ek(d, k, eta1):
    """FIPS 203 Algorithm 13 (K-PKE.KeyGen), encapsulation key."""
    g = hashlib.sha3_512(d + bytes([k])).digest()
    rho, sigma = g[:32], g[32:]
    s = [cbd(sigma, n, eta1) for n in range(k)]
    e = [cbd(sigma, k + n, eta1) for n in range(k)]
    sh, eh = [ntt(x) for x in s], [ntt(x) for x in e]
    ek = bytearray()
    for i in range(k):
        t = eh[i]
        for j in range(k):
            t = [(x + y) % Q for x, y in zip(t, base_mul(sample_ntt(rho, j, i), sh[j]))]
        for z in range(0, 256, 2):  # ByteEncode_12
            ek += bytes([t[z] & 0xff, (t[z] >> 8) | ((t[z + 1] & 15) << 4), t[z + 1] >> 4])
    return bytes(ek) + rho


KEMS = {"ml-kem-512": (2, 3, 800), "ml-kem-768": (3, 2, 1184), "ml-kem-1024": (4, 2, 1568)}


def test_keccak_matches_hashlib():
    rng = random.Random(5)
    for n in (0, 1, 71, 72, 73, 135, 136, 137, 167, 168, 169, 500):
        m = bytes(rng.randrange(256) for _ in range(n))
        assert IA.keccak(0, m, 32) == hashlib.sha3_256(m).digest()
        assert IA.keccak(1, m, 64) == hashlib.sha3_512(m).digest()
        for out in (1, 168, 169, 600):
            assert IA.keccak(2, m, out) == hashlib.shake_128(m).digest(out)
            assert IA.keccak(3, m, out) == hashlib.shake_256(m).digest(out)


def test_ntt_is_the_remainder_definition():
    # NTT(f)[2i] + NTT(f)[2i+1] X = f mod (X^2 - gamma_i): reduce f by
    # X^2 -> gamma_i, i.e. even coefficients weighted by gamma^(m), odd likewise
    rng = random.Random(6)
    f = [rng.randrange(Q) for _ in range(256)]
    fh = ntt(f)
    for i in range(128):
        g = GAMMA[i]
        r0 = sum(f[2 * m] * pow(g, m, Q) for m in range(128)) % Q
        r1 = sum(f[2 * m + 1] * pow(g, m, Q) for m in range(128)) % Q
        assert (fh[2 * i], fh[2 * i + 1]) == (r0, r1), i
    # base-case products are the products in Z_q[X]/(X^256 + 1)
    h = [rng.randrange(Q) for _ in range(256)]
    prod = [0] * 256
    for a in range(256):
        for b in range(256):
            if a + b < 256:
                prod[a + b] = (prod[a + b] + f[a] * h[b]) % Q
            else:
                prod[a + b - 256] = (prod[a + b - 256] - f[a] * h[b]) % Q
    assert base_mul(fh, ntt(h)) == ntt(prod)


@pytest.mark.parametrize("kem", sorted(KEMS))
def test_mlkem_keygen_matches_fips203(kem):
    k, eta1, n = KEMS[kem]
    rng = random.Random(len(kem))
    for _ in range(3):
        seed = bytes(rng.randrange(256) for _ in range(64))
        ek = IA.mlkem_keygen(kem, seed)
        assert len(ek) == n
        assert ek == kpke_ek(seed[:32], k, eta1)
        assert ek[-32:] == hashlib.sha3_512(seed[:32] + bytes([k])).digest()[:32]  # rho
    # z (the second half of the seed) enters only the decapsulation key
    assert IA.mlkem_keygen(kem, seed[:32] + bytes(32)) == IA.mlkem_keygen(kem, seed)


def pcg32_seed(state):
    """rand_core SeedableRng::seed_from_u64: PCG32 fills the 32-byte seed."""
    mul, inc, m64 = 6364136223846793005, 11634580027462260723, (1 << 64) - 1
    out = bytearray()
    for _ in range(8):
        state = (state * mul + inc) & m64
        xs = (((state >> 18) ^ state) >> 27) & 0xffffffff
        rot = state >> 59
        out += struct.pack("<I", ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xffffffff)
    return bytes(out)


@pytest.mark.parametrize("seed,index", [(7, 0), (7, 1), (11, 999_999), (19, 31_999_999), (1, 2**33 + 5)])
def test_record_seed_is_chacha_block_index(seed, index):
    key = pcg32_seed(seed)
    words = O.chacha20_block(key, [index & 0xffffffff, index >> 32, 0, 0])
    assert IA.dataset_record_seed(seed, index) == np.asarray(words, dtype="<u4").tobytes()


def test_generated_records_and_pqkeyv2_file(tmp_path):
    kem, count, seed = "ml-kem-768", 300, 7
    recs = IA.generate_records(kem, count, seed=seed)
    for i in (0, 1, 137, 299):
        ident, pk = recs[i]
        assert pk == IA.mlkem_keygen(kem, IA.dataset_record_seed(seed, i))
        assert ident == hashlib.sha3_256(b"pq-key:v1:" + kem.encode() + pk).digest()
    assert len({r[0] for r in recs}) == count
    # a slice of the stream equals the same records generated from the start
    assert IA.generate_records(kem, 5, seed=seed, start=135) == recs[135:140]
    path = str(tmp_path / "ds.bin")
    IA.generate_dataset(path, kem, count, seed=seed, threads=3)
    blob = open(path, "rb").read()
    assert blob[:8] == b"PQKEYV2\0" and blob[8] == len(kem) and blob[9:9 + len(kem)] == kem.encode()
    o = 9 + len(kem)
    assert struct.unpack_from("<IH", blob, o) == (count, 1184)
    o += 6
    assert len(blob) == o + count * (32 + 1184)
    for i in range(count):
        r = blob[o + i * 1216:o + (i + 1) * 1216]
        assert (r[:32], r[32:]) == recs[i], i


def test_generator_errors(tmp_path):
    with pytest.raises(IA.InspireError, match="unknown KEM"):
        IA.generate_dataset(str(tmp_path / "x.bin"), "ml-kem-999", 10)
    with pytest.raises(IA.InspireError, match="greater than zero"):
        IA.generate_dataset(str(tmp_path / "x.bin"), "ml-kem-768", 0)
