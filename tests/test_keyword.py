"""Keyword layer on the host: SipHash-2-4 and the cuckoo positions
(pir/keyword/cuckoo.rs:98-132), in the product library and in the oracle.
SipHash-2-4 is pinned by the test vectors of its paper (Aumasson and
Bernstein 2012, appendix: key 00..0f, message 00..len-1)."""
import os

import pytest

import inspire_amd as IA
import oracle as O

K0 = int.from_bytes(bytes(range(8)), "little")
K1 = int.from_bytes(bytes(range(8, 16)), "little")
VECTORS = {0: 0x726fdb47dd0e0e31, 1: 0x74f839c593dc67fd, 15: 0xa129ca6149be45e5, 63: 0x958a324ceb064572}


@pytest.mark.parametrize("length", sorted(VECTORS))
def test_siphash24_paper_vectors(length):
    msg = bytes(range(length))
    assert O.siphash24(K0, K1, msg) == VECTORS[length]
    assert IA.siphash24(K0, K1, msg) == VECTORS[length]


def test_cuckoo_positions_library_equals_oracle():
    for n in (32768, 2097152, 1000003):
        for _ in range(50):
            key = os.urandom(32)
            assert IA.cuckoo_positions(key, n) == O.cuckoo_positions(key, n)
    # the hash index enters both SipHash keys (cuckoo.rs:116-118)
    key = bytes(32)
    p = O.cuckoo_positions(key, 1 << 40)
    k0 = int.from_bytes(O.CUCKOO_SEED[:8], "little")
    k1 = int.from_bytes(O.CUCKOO_SEED[8:], "little")
    assert p == [O.siphash24(k0, k1, key) % (1 << 40), O.siphash24(k0 ^ 1, k1 ^ (1 << 32), key) % (1 << 40)]


def test_public_key_id_is_sha3_of_domain_kem_and_key():
    import hashlib
    pk = bytes(range(256)) * 4
    assert O.public_key_id(pk) == hashlib.sha3_256(b"pq-key:v1:ml-kem-768" + pk).digest()
