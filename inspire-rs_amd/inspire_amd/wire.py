"""TCP framing and the handshake of the keyword-PIR protocol (src/protocol/wire.rs).

A frame is [8 bytes: payload length, big-endian u64] [1 byte: message type]
[payload] (send_msg / recv_msg, wire.rs:41-86).  Payloads over 64 MiB are
refused before they are read.  Message types: MSG_HANDSHAKE (the server's
PublicParams as JSON), MSG_KEYWORD_QUERY (serialize_keyword_query bytes) and
MSG_KEYWORD_RESPONSE (serialize_keyword_response bytes), wire.rs:15-17.

The handshake JSON has serde_json's layout of PublicParams and
KeywordPirHandshake (wire.rs:19-36): fields in declaration order, compact
separators, the cuckoo seed as an array of 16 numbers."""
import json
import socket
import struct

MSG_HANDSHAKE = 0x01
MSG_KEYWORD_QUERY = 0x05
MSG_KEYWORD_RESPONSE = 0x06
MAX_MSG_SIZE = 64 << 20
PAYLOAD_TIMEOUT_S = 30.0


class WireError(IOError):
    pass


def send_msg(sock, msg_type, payload):
    """One frame in a single write (wire.rs:41-51)."""
    payload = bytes(payload)
    sock.sendall(struct.pack(">Q", len(payload)) + bytes([msg_type & 0xFF]) + payload)


def _read_exact(sock, n):
    chunks, got = [], 0
    while got < n:
        b = sock.recv(min(n - got, 1 << 20))
        if not b:
            raise EOFError("connection closed after %d of %d bytes" % (got, n))
        chunks.append(b)
        got += len(b)
    return b"".join(chunks)


def recv_msg(sock):
    """(msg_type, payload) (wire.rs:56-86).  The length is read under the
    socket's own timeout; once it has arrived the rest may take up to 30 s.
    EOFError when the peer closed before a frame began."""
    head = _read_exact(sock, 8)
    (length,) = struct.unpack(">Q", head)
    if length > MAX_MSG_SIZE:
        raise WireError("message too large: %d bytes" % length)
    prev = sock.gettimeout()
    sock.settimeout(PAYLOAD_TIMEOUT_S)
    try:
        kind = _read_exact(sock, 1)[0]
        payload = _read_exact(sock, length)
    finally:
        sock.settimeout(prev)
    return kind, payload


def public_params_to_json(pp):
    """serde_json::to_vec(&PublicParams) of a public_params dict."""
    obj = {
        "num_items": int(pp["num_items"]),
        "item_size_bits": int(pp["item_size_bits"]),
        "dim0": int(pp["dim0"]),
        "keyword_pir": {
            "cuckoo_seed": list(bytes(pp["cuckoo_seed"])),
            "num_hashes": int(pp["num_hashes"]),
            "key_size": int(pp["key_size"]),
            "value_size": int(pp["value_size"]),
            "entry_size": int(pp["entry_size"]),
            "num_records": int(pp.get("num_records", 0)),
        },
    }
    return json.dumps(obj, separators=(",", ":")).encode()


def public_params_from_json(data):
    """The public_params dict of a handshake; ValueError("server is not in
    keyword PIR mode") without a keyword_pir object (client/keyword.rs:64-67)."""
    obj = json.loads(bytes(data).decode())
    kw = obj.get("keyword_pir")
    if not kw:
        raise ValueError("server is not in keyword PIR mode")
    seed = bytes(kw["cuckoo_seed"])
    if len(seed) != 16:
        raise ValueError("cuckoo seed must be 16 bytes, got %d" % len(seed))
    return {"num_items": int(obj["num_items"]), "item_size_bits": int(obj["item_size_bits"]),
            "dim0": int(obj["dim0"]), "cuckoo_seed": seed, "num_hashes": int(kw["num_hashes"]),
            "key_size": int(kw["key_size"]), "value_size": int(kw["value_size"]),
            "entry_size": int(kw["entry_size"]), "num_records": int(kw.get("num_records", 0))}


def connect(address, timeout=None):
    """A TCP connection with TCP_NODELAY, as the reference's client opens it."""
    host, port = address.rsplit(":", 1)
    s = socket.create_connection((host, int(port)), timeout=timeout)
    s.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
    return s
