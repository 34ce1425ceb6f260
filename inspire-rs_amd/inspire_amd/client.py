"""Keyword-PIR TCP client (src/bin/client.rs).

python -m inspire_amd.client [--server 127.0.0.1:8082] --id HEX [--hamming 0]

Connects, reads the handshake (the server's public params), builds the
library's KeywordClient from it, sends one keyword query for the id (key_size
bytes, 32 for public-key ids) and
prints the public key it extracts from the response, or "not found".
--hamming 0 is the reference client's zero secret; a positive weight draws a
ternary secret instead."""
import argparse
import sys
import time

from . import KeywordClient
from . import wire


def fetch(address, key, hamming=0, rng_seed=0, timeout=60.0):
    """One query over TCP: (value or None, elapsed seconds)."""
    with wire.connect(address, timeout=timeout) as s:
        kind, hs = wire.recv_msg(s)
        if kind != wire.MSG_HANDSHAKE:
            raise wire.WireError("expected handshake, got message type %d" % kind)
        pp = wire.public_params_from_json(hs)
        if len(key) != pp["key_size"]:
            raise ValueError("expected %d bytes for --id, got %d" % (pp["key_size"], len(key)))
        client = KeywordClient.setup_from_public_params(pp, hamming=hamming, rng_seed=rng_seed)
        t0 = time.perf_counter()
        query = client.query(key)
        wire.send_msg(s, wire.MSG_KEYWORD_QUERY, client.serialize_query(query))
        kind, resp = wire.recv_msg(s)
        if kind != wire.MSG_KEYWORD_RESPONSE:
            raise wire.WireError("expected keyword response, got %d" % kind)
        value = client.extract_from_serialized_response(key, query, resp)
        return value, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser(description="Keyword PIR client for ML-KEM public keys")
    ap.add_argument("--server", default="127.0.0.1:8082")
    ap.add_argument("--id", required=True, help="32-byte public key id as hex")
    ap.add_argument("--hamming", type=int, default=0, help="secret key weight (0: the reference's zero secret)")
    a = ap.parse_args(argv)
    hexid = a.id[2:] if a.id.startswith("0x") else a.id
    key = bytes.fromhex(hexid)
    try:
        value, elapsed = fetch(a.server, key, hamming=a.hamming)
    except ValueError as e:
        print(e, file=sys.stderr)
        return 1
    print("query id: %s" % key.hex())
    print("public key: %s" % value.hex() if value is not None else "not found")
    print("elapsed: %.1fms" % (elapsed * 1e3))
    return 0


if __name__ == "__main__":
    sys.exit(main())
