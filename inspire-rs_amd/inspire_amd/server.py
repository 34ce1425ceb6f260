"""Keyword-PIR TCP server (src/bin/server.rs) over the library's batching engine.

python -m inspire_amd.server --dataset FILE [--listen 0.0.0.0:8082]
    [--buckets N] [--dim0 D] [--estimate-memory] [--max-batch 64] [--window-ms 0]

As the reference's server: load the PQKEYV2 dataset, print the GPU memory
check, build the server (cuckoo table, DB encode and offline precomputation on
the device), then per connection send the handshake (MSG_HANDSHAKE, the public
params as JSON) and answer MSG_KEYWORD_QUERY frames with MSG_KEYWORD_RESPONSE
frames until the peer closes.  Unlike the reference, whose accept loop serves
one connection at a time and one query at a time (server.rs:116-158), every
connection runs on its own thread and submits to one continuous-batching
engine, so queries from all connections share batched passes over the
database and the key-switch material."""
import argparse
import socket
import sys
import threading
import time

from . import Engine, InspireError, lib, plan_shape, setup_keyword_server
from . import wire


def _gib(n):
    return "%.2f GiB" % (n / (1024.0 ** 3))


def _log(msg):
    print("[init %6ds] %s" % (int(time.time()) % 100000, msg), file=sys.stderr, flush=True)


def dataset_header(path):
    """(kem name, record count, public key length) of a PQKEYV2 file (dataset.rs:95-142)."""
    with open(path, "rb") as f:
        head = f.read(9)
        if head[:8] != b"PQKEYV2\0":
            raise ValueError("not a PQKEYV2 dataset: %s" % path)
        name = f.read(head[8]).decode()
        count = int.from_bytes(f.read(4), "little")
        pk_len = int.from_bytes(f.read(2), "little")
    return name, count, pk_len


class KeywordTcpServer:
    """Serves one keyword Server over TCP through an Engine.  start() binds and
    returns the address; stop() closes the listener, the connections and the
    engine (not the Server)."""

    def __init__(self, server, public_params, listen="127.0.0.1:0", max_batch=64, window_ms=0.0, log=None):
        self.server, self.pp = server, public_params
        self.listen = listen
        self.engine = Engine(server, nq=public_params["num_hashes"], max_batch=max_batch, window_ms=window_ms)
        self.handshake = wire.public_params_to_json(public_params)
        self.log = log or (lambda m: None)
        self._sock = None
        self._threads = []
        self._conns = set()
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self.answered = 0

    def start(self):
        host, port = self.listen.rsplit(":", 1)
        s = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        s.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        s.bind((host, int(port)))
        s.listen(64)
        s.settimeout(0.2)
        self._sock = s
        t = threading.Thread(target=self._accept_loop, daemon=True)
        t.start()
        self._threads.append(t)
        h, p = s.getsockname()[:2]
        return "%s:%d" % (h, p)

    def _accept_loop(self):
        while not self._stop.is_set():
            try:
                conn, _ = self._sock.accept()
            except socket.timeout:
                continue
            except OSError:
                break
            conn.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
            with self._lock:
                self._conns.add(conn)
            t = threading.Thread(target=self._serve, args=(conn,), daemon=True)
            t.start()
            self._threads.append(t)

    def _serve(self, conn):
        try:
            wire.send_msg(conn, wire.MSG_HANDSHAKE, self.handshake)
            while not self._stop.is_set():
                try:
                    kind, query = wire.recv_msg(conn)
                except EOFError:
                    break
                start = time.perf_counter()
                if kind != wire.MSG_KEYWORD_QUERY:
                    self.log("client error: unexpected message type %d" % kind)
                    break
                ticket = self.engine.submit(query)
                resp, _ = self.engine.wait(ticket)
                wire.send_msg(conn, wire.MSG_KEYWORD_RESPONSE, resp)
                with self._lock:
                    self.answered += 1
                self.log("query answered in %.1fms" % ((time.perf_counter() - start) * 1e3))
        except (OSError, wire.WireError, InspireError) as e:
            if not self._stop.is_set():
                self.log("client error: %s" % e)
        finally:
            with self._lock:
                self._conns.discard(conn)
            conn.close()

    def stop(self):
        self._stop.set()
        if self._sock:
            self._sock.close()
        with self._lock:
            conns = list(self._conns)
        for c in conns:
            try:
                c.shutdown(socket.SHUT_RDWR)
            except OSError:
                pass
        for t in self._threads:
            t.join(timeout=30)
        self.engine.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description="Static keyword PIR server for ML-KEM public keys")
    ap.add_argument("--dataset", required=True, help="input dataset file generated by gen_dataset")
    ap.add_argument("--listen", default="0.0.0.0:8082")
    ap.add_argument("--buckets", type=int, default=None, help="number of cuckoo buckets / PIR items")
    ap.add_argument("--dim0", type=int, default=None, help="first-dimension size override")
    ap.add_argument("--estimate-memory", action="store_true", help="print the GPU memory estimate and exit")
    ap.add_argument("--max-batch", type=int, default=64, help="largest batch of the engine")
    ap.add_argument("--window-ms", type=float, default=0.0, help="arrival window before a partial batch runs")
    a = ap.parse_args(argv)

    _log("loading dataset")
    kem, count, pk_len = dataset_header(a.dataset)
    if count == 0:
        print("dataset must contain at least one record", file=sys.stderr)
        return 1
    _log("loaded %d %s records (public key %d bytes)" % (count, kem, pk_len))
    _log("checking GPU memory requirements")
    plan = plan_shape(count, 32, pk_len, buckets=a.buckets, dim0=a.dim0)
    avail = int(lib().gpu_available_memory_bytes())
    print("startup GPU memory check: require %s (%d bytes), available %s (%d bytes)" % (
        _gib(plan["peak_bytes"]), plan["peak_bytes"], _gib(avail), avail))
    print("estimate details: buckets %d, dim0 %d, db_rows %d, db_cols %d, instances %d" % (
        plan["num_items"], plan["dim0"], plan["db_rows"], plan["db_cols"], plan["instances"]))
    if a.estimate_memory:
        return 0
    if avail and plan["peak_bytes"] > avail:
        print("insufficient GPU memory for keyword PIR setup", file=sys.stderr)
        return 1
    _log("initializing keyword PIR runtime")
    srv, pp = setup_keyword_server(a.dataset, a.buckets, a.dim0)
    _log("keyword PIR runtime ready")
    tcp = KeywordTcpServer(srv, pp, listen=a.listen, max_batch=a.max_batch, window_ms=a.window_ms,
                           log=lambda m: print(m, flush=True))
    addr = tcp.start()
    print("serving %d %s public keys on %s" % (count, kem, addr))
    print("entry size %d bytes, buckets %d, dim0 %d" % (pp["item_size_bits"] // 8, pp["num_items"], pp["dim0"]),
          flush=True)
    try:
        while True:
            time.sleep(3600)
    except KeyboardInterrupt:
        pass
    finally:
        tcp.stop()
        srv.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
