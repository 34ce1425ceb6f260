"""Python binding of libinspire_hip.so (the C ABI in include/inspire_hip.h).

The answer path of the inspire-rs keyword-PIR server
(src/pir/server/keyword.rs:534-628) runs in hand-written gfx950 kernels; this
module only loads the library and marshals pointers.  Device buffers are torch
tensors (plumbing); every computation happens in the HIP library.  There is no
CPU fallback: loading fails loudly when the library has not been built.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libinspire_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "inspire_hip.h")

N = 2048
TEXP = 3
NINNER = 3069
RESP_BYTES = 12288

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_CAPACITY = 0, -1, -2, -3, -4, -5

# Synthetic generator kinds and stream ids (shared with the oracle).
KIND_CONDENSED, KIND_MODQ, KIND_U16, KIND_CRT, KIND_SMALL = 0, 1, 2, 3, 4
STREAM_DB, STREAM_BT, STREAM_BTB, STREAM_BTH, STREAM_AHAT = 1, 2, 3, 4, 5
STREAM_Y, STREAM_Z, STREAM_ROW, STREAM_BODY = 16, 17, 18, 19

_lib = None


class PublicParamsT(ctypes.Structure):
    _fields_ = [("num_items", ctypes.c_int64), ("item_size_bits", ctypes.c_int64), ("dim0", ctypes.c_int64),
                ("cuckoo_seed", ctypes.c_uint8 * 16), ("num_hashes", ctypes.c_int32), ("key_size", ctypes.c_int64),
                ("value_size", ctypes.c_int64), ("entry_size", ctypes.c_int64), ("num_records", ctypes.c_int64)]


class TuneResult(ctypes.Structure):
    """inspire_tune_t (include/inspire_hip.h)."""
    _fields_ = [("packing_variant", ctypes.c_int), ("gemv_variant", ctypes.c_int), ("fold_variant", ctypes.c_int),
                ("packing_variants", ctypes.c_int), ("gemv_variants", ctypes.c_int),
                ("packing_ms", ctypes.c_double * 16), ("gemv_ms", ctypes.c_double * 8), ("fold_ms", ctypes.c_double * 3),
                ("nbatch", ctypes.c_int), ("batch_sizes", ctypes.c_int * 16), ("batch_qps", ctypes.c_double * 16),
                ("suggested_batch", ctypes.c_int)]


class InspireError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("inspire error %d: %s" % (code, msg))
        self.code = code


class PlanT(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("num_items", "item_size_bits", "dim0", "db_rows", "db_cols",
                                               "instances", "interpolate_degree", "c")] + \
               [(n, ctypes.c_uint64) for n in ("db_bytes", "hint_bytes", "prep_pack_bytes",
                                               "r_buffers_bytes", "bold_t_bytes", "peak_bytes")]


class ShapeT(ctypes.Structure):
    _fields_ = [("db_rows", ctypes.c_int64), ("instances", ctypes.c_int64),
                ("interpolate_degree", ctypes.c_int32), ("max_queries", ctypes.c_int32)]


def lib():
    """Load libinspire_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libinspire_hip.so not built: run __graft_entry__.build() "
                               "(expected at %s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, u64, u32, sz = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_uint32, ctypes.c_size_t)
        L.inspire_last_error.restype = ctypes.c_char_p
        L.inspire_plan_shape.argtypes = [i64, i64, i64, i64, i64, ctypes.POINTER(PlanT)]
        L.inspire_server_create.argtypes = [ctypes.POINTER(ShapeT), ctypes.POINTER(vp)]
        L.inspire_server_create_lean.argtypes = [ctypes.POINTER(ShapeT), ctypes.POINTER(vp)]
        L.inspire_server_destroy.argtypes = [vp]
        L.inspire_server_destroy.restype = None
        L.inspire_server_set_db.argtypes = [vp, vp, i32]
        L.inspire_server_encode_db.argtypes = [vp, vp, i32]
        L.inspire_encode_db_device.argtypes = [vp, vp, i64, i64, i32, vp]
        L.inspire_server_set_precomp.argtypes = [vp, vp, vp, vp, vp, i32]
        L.inspire_server_set_stash.argtypes = [vp, vp, u32, u32]
        L.inspire_server_setup_offline.argtypes = [vp]
        L.inspire_server_swap_state.argtypes = [vp, vp]
        L.inspire_engine_swap_state.argtypes = [vp, vp]
        L.inspire_setup_keyword_server.argtypes = [ctypes.c_char_p, i64, i64, ctypes.POINTER(vp),
                                                   ctypes.POINTER(PublicParamsT)]
        L.inspire_setup_keyword_records.argtypes = [vp, u32, vp, u32, i64, i64, i64, ctypes.POINTER(vp),
                                                    ctypes.POINTER(PublicParamsT)]
        L.inspire_cuckoo_positions.argtypes = [vp, sz, i64, ctypes.POINTER(i64)]
        L.inspire_siphash24.argtypes = [u64, u64, vp, sz]
        L.inspire_siphash24.restype = u64
        L.inspire_setup_offline_device.argtypes = [vp, i64, i64, vp, vp, vp, vp, vp]
        L.inspire_server_fill_synthetic.argtypes = [vp, u64]
        L.inspire_server_response_bytes.argtypes = [vp, i32]
        L.inspire_server_response_bytes.restype = sz
        L.inspire_respond_serialized.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.inspire_respond_device.argtypes = [vp, vp, sz, i32, vp, vp]
        L.inspire_respond_device_noalloc.argtypes = [vp, vp, sz, i32, vp, vp]
        L.inspire_noalloc_scratch_bytes.argtypes = [ctypes.POINTER(ShapeT), i32]
        L.inspire_noalloc_scratch_bytes.restype = sz
        L.inspire_server_profile_enable.argtypes = [vp, i32]
        L.inspire_server_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64), i32]
        L.inspire_server_stream.argtypes = [vp]
        L.inspire_server_stream.restype = vp
        L.inspire_synth_fill.argtypes = [vp, u64, u32, u64, u64, i32, vp]
        L.inspire_unpack_query_device.argtypes = [vp, i32, i64, vp, vp, vp, vp, vp]
        L.inspire_gemv_device.argtypes = [vp, vp, i32, vp, i64, i64, vp]
        L.inspire_gemv_mfma_device.argtypes = [vp, vp, i32, vp, i64, i64, vp, vp]
        L.inspire_gemv_mfma_scratch_bytes.argtypes = [i64]
        L.inspire_gemv_mfma_scratch_bytes.restype = sz
        L.inspire_server_set_gemv_kernel.argtypes = [vp, i32]
        L.inspire_server_set_slots.argtypes = [vp, i32]
        L.inspire_server_set_packing_kernel.argtypes = [vp, i32]
        L.inspire_server_set_max_batch.argtypes = [vp, i32]
        L.inspire_server_set_batch_kernels.argtypes = [vp, i32, i32, i32]
        L.inspire_server_set_graphs.argtypes = [vp, i32]
        L.inspire_server_autotune.argtypes = [vp, i32, i32, i32, ctypes.POINTER(TuneResult)]
        L.inspire_respond_batch_device.argtypes = [vp, vp, sz, sz, i32, i32, vp, sz, vp]
        L.inspire_packing_device.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp]
        L.inspire_post_device.argtypes = [vp, vp, i64, vp]
        L.inspire_fold_device.argtypes = [vp, vp, vp, vp, vp, i64, i32, i32, vp]
        L.inspire_engine_create.argtypes = [vp, i32, i32, ctypes.c_double, i32, ctypes.POINTER(vp)]
        L.inspire_engine_submit.argtypes = [vp, vp, sz, ctypes.POINTER(u64)]
        L.inspire_engine_wait.argtypes = [vp, u64, vp, sz, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_double)]
        L.inspire_engine_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i32)]
        L.inspire_engine_destroy.argtypes = [vp]
        L.inspire_engine_destroy.restype = None
        L.inspire_dataset_generate.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i64, u64, i32]
        L.inspire_dataset_generate_records.argtypes = [ctypes.c_char_p, u64, i64, i64, vp, vp, i32]
        L.inspire_dataset_record_seed.argtypes = [u64, u64, vp]
        L.inspire_mlkem_keygen.argtypes = [ctypes.c_char_p, vp, vp, sz]
        L.inspire_keccak.argtypes = [i32, vp, sz, vp, sz]
        L.inspire_client_new.argtypes = [ctypes.POINTER(PublicParamsT), i32, u64, ctypes.POINTER(vp)]
        L.inspire_client_free.argtypes = [vp]
        L.inspire_client_free.restype = None
        L.inspire_client_secret.argtypes = [vp, vp]
        L.inspire_client_positions.argtypes = [vp, vp, sz, ctypes.POINTER(i64)]
        L.inspire_client_query_bytes.argtypes = [vp]
        L.inspire_client_query_bytes.restype = sz
        L.inspire_client_query.argtypes = [vp, vp, sz, ctypes.POINTER(i64), vp, sz, ctypes.POINTER(sz)]
        L.inspire_client_extract.argtypes = [vp, vp, sz, ctypes.POINTER(i64), vp, sz, vp, sz, ctypes.POINTER(sz),
                                             ctypes.POINTER(i32)]
        L.gpu_available_memory_bytes.restype = u64
        L.gpu_packing_online_get_d_y_all.restype = vp
        L.gpu_packing_online_get_d_y_bar_all.restype = vp
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise InspireError(rc, lib().inspire_last_error().decode())
    return rc


def plan_shape(record_count, key_size, value_size, buckets=None, dim0=None):
    """KeywordServerSetupPlan::from_shape + memory estimate (keyword.rs:104-233)."""
    p = PlanT()
    check(lib().inspire_plan_shape(record_count, key_size, value_size, buckets or 0, dim0 or 0,
                                   ctypes.byref(p)))
    return {f: getattr(p, f) for f, _ in PlanT._fields_}


def _ptr(t):
    """Device or host address of a torch tensor / numpy array / int."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


class Server:
    """KeywordServer answer path resident on one GPU (keyword.rs:49-63, 534-640)."""

    def __init__(self, db_rows, instances, interpolate_degree, max_queries=2, lean=False):
        self.db_rows, self.instances, self.d = db_rows, instances, interpolate_degree
        self.c = instances // interpolate_degree
        shape = ShapeT(db_rows, instances, interpolate_degree, max_queries)
        h = ctypes.c_void_p()
        create = lib().inspire_server_create_lean if lean else lib().inspire_server_create
        check(create(ctypes.byref(shape), ctypes.byref(h)))
        self.h = h

    @classmethod
    def _adopt(cls, h, plan):
        self = cls.__new__(cls)
        self.db_rows, self.instances, self.d = plan["db_rows"], plan["instances"], plan["interpolate_degree"]
        self.c = self.instances // self.d
        self.h = h
        return self

    def close(self):
        if self.h:
            lib().inspire_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill_synthetic(self, seed):
        check(lib().inspire_server_fill_synthetic(self.h, seed))

    def set_db(self, db, host=True):
        check(lib().inspire_server_set_db(self.h, _ptr(db), 1 if host else 0))

    def set_graphs(self, on):
        check(lib().inspire_server_set_graphs(self.h, 1 if on else 0))

    def set_batch_kernels(self, packing_variant, gemv_variant, fold_variant):
        check(lib().inspire_server_set_batch_kernels(self.h, packing_variant, gemv_variant, fold_variant))

    def autotune(self, count, nq=2, reps=3):
        """Times the batched kernel variants on this device, keeps the fastest
        and sweeps the batch size; returns the measurements as a dict."""
        r = TuneResult()
        check(lib().inspire_server_autotune(self.h, count, nq, reps, ctypes.byref(r)))
        return {"packing_variant": r.packing_variant, "gemv_variant": r.gemv_variant,
                "fold_variant": r.fold_variant, "packing_ms": list(r.packing_ms[:r.packing_variants]),
                "gemv_ms": list(r.gemv_ms[:r.gemv_variants]), "fold_ms": list(r.fold_ms[1:3]),
                "batch_qps": {int(r.batch_sizes[i]): r.batch_qps[i] for i in range(r.nbatch)},
                "suggested_batch": r.suggested_batch}

    def encode_db(self, raw, host=True):
        """Raw column-major DB in, encoded on the device into the server's DB
        (keyword.rs:404-424)."""
        check(lib().inspire_server_encode_db(self.h, _ptr(raw), 1 if host else 0))

    def set_precomp(self, bt, btb, bth, a_hat, host=True):
        check(lib().inspire_server_set_precomp(self.h, _ptr(bt), _ptr(btb), _ptr(bth), _ptr(a_hat),
                                               1 if host else 0))

    def swap_state(self, staged):
        """Live DB update: exchange resident state with a server of the same shape."""
        check(lib().inspire_server_swap_state(self.h, staged.h))

    def setup_offline(self):
        """KeywordServer::setup's offline precomputation on the device from the
        server's DB (keyword.rs:427-458)."""
        check(lib().inspire_server_setup_offline(self.h))

    def set_stash(self, entries):
        entries = list(entries)
        size = len(entries[0]) if entries else 0
        blob = b"".join(entries)
        buf = ctypes.create_string_buffer(blob, len(blob) or 1)
        check(lib().inspire_server_set_stash(self.h, ctypes.cast(buf, ctypes.c_void_p), size, len(entries)))

    def stream(self):
        return lib().inspire_server_stream(self.h)

    def response_bytes(self, nq):
        """Serialized response size for a query with nq hash queries (stash included)."""
        return lib().inspire_server_response_bytes(self.h, nq)

    def respond_serialized(self, query):
        """respond_to_serialized_query (keyword.rs:636-640): bytes in, bytes out."""
        nq = query[0] if query else 0
        cap = lib().inspire_server_response_bytes(self.h, nq)
        out = ctypes.create_string_buffer(max(cap, 1))
        qb = ctypes.create_string_buffer(bytes(query), len(query) or 1)
        n = ctypes.c_size_t()
        check(lib().inspire_respond_serialized(self.h, ctypes.cast(qb, ctypes.c_void_p), len(query),
                                               ctypes.cast(out, ctypes.c_void_p), cap, ctypes.byref(n)))
        return out.raw[:n.value]

    def set_slots(self, count):
        check(lib().inspire_server_set_slots(self.h, count))

    def set_packing_kernel(self, kind):
        check(lib().inspire_server_set_packing_kernel(self.h, kind))

    def set_max_batch(self, max_batch):
        check(lib().inspire_server_set_max_batch(self.h, max_batch))

    def respond_batch_device(self, d_queries, q_stride, length, nq, count, d_out, out_stride, stream=None):
        check(lib().inspire_respond_batch_device(self.h, _ptr(d_queries), q_stride, length, nq, count, _ptr(d_out),
                                                 out_stride, stream))

    def set_gemv_kernel(self, kind):
        check(lib().inspire_server_set_gemv_kernel(self.h, kind))

    def profile(self, enable=True):
        check(lib().inspire_server_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self, reset=True):
        """Summed ms per stage [unpack, packing, post, gemv, fold] and the respond count."""
        ms = (ctypes.c_double * 5)()
        cnt = ctypes.c_int64()
        check(lib().inspire_server_profile_read(self.h, ms, ctypes.byref(cnt), 1 if reset else 0))
        return list(ms), cnt.value

    def respond_device(self, d_query, length, nq, d_out, stream=None):
        check(lib().inspire_respond_device(self.h, _ptr(d_query), length, nq, _ptr(d_out), stream))

    def respond_device_noalloc(self, d_query, length, nq, d_out, stream=None):
        """respond_device with the working buffers in the library's per-GPU
        scratch array (no device allocation; calls on one GPU must not overlap)."""
        check(lib().inspire_respond_device_noalloc(self.h, _ptr(d_query), length, nq, _ptr(d_out), stream))


class Engine:
    """Continuous-batching front end over a Server (inspire_engine_* in the C ABI).

    submit() may be called from any thread and returns a ticket; wait(ticket)
    blocks until that query's batch is done and returns (response bytes,
    submit-to-ready latency in ms).  Every ticket is waited exactly once."""

    def __init__(self, server, nq=2, max_batch=64, window_ms=0.0, capacity=1024):
        self.server = server
        h = ctypes.c_void_p()
        check(lib().inspire_engine_create(server.h, nq, max_batch, float(window_ms), capacity, ctypes.byref(h)))
        self.h = h
        self.nq = nq
        self._resp_cap = server.response_bytes(nq)

    def submit(self, query):
        q = bytes(query)
        t = ctypes.c_uint64()
        check(lib().inspire_engine_submit(self.h, q, len(q), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        cap = self.server.response_bytes(self.nq)
        n = ctypes.c_size_t()
        lat = ctypes.c_double()
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = lib().inspire_engine_wait(self.h, ticket, buf, cap, ctypes.byref(n), ctypes.byref(lat))
            if rc == ERR_CAPACITY and n.value > cap:
                cap = n.value  # an answer framed with a larger (earlier) stash
                continue
            check(rc)
            return buf.raw[:n.value], lat.value

    def swap_state(self, staged):
        """Live DB update under the engine (inspire_engine_swap_state)."""
        check(lib().inspire_engine_swap_state(self.h, staged.h))

    def stats(self):
        b, q, m = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check(lib().inspire_engine_stats(self.h, ctypes.byref(b), ctypes.byref(q), ctypes.byref(m)))
        return {"batches": b.value, "queries": q.value, "max_batch": m.value}

    def close(self):
        if self.h:
            lib().inspire_engine_destroy(self.h)
            self.h = None


NOALLOC_SCRATCH_BYTES = 96 << 20


def noalloc_scratch_bytes(db_rows, instances, interpolate_degree, nq, max_queries=2):
    """Scratch bytes inspire_respond_device_noalloc needs for a shape (host only)."""
    return lib().inspire_noalloc_scratch_bytes(ctypes.byref(ShapeT(db_rows, instances, interpolate_degree,
                                                                   max_queries)), nq)


def synth_fill(d_out, seed, stream_id, count, kind, start=0, stream=None):
    check(lib().inspire_synth_fill(_ptr(d_out), seed, stream_id, start, count, kind, stream))


def _public_params(pp):
    return {"num_items": pp.num_items, "item_size_bits": pp.item_size_bits, "dim0": pp.dim0,
            "cuckoo_seed": bytes(pp.cuckoo_seed), "num_hashes": pp.num_hashes, "key_size": pp.key_size,
            "value_size": pp.value_size, "entry_size": pp.entry_size, "num_records": pp.num_records}


def setup_keyword_server(path, buckets=None, dim0=None):
    """setup_keyword_server (dataset/keyword.rs:29-44) from a PQKEYV2 file:
    returns (Server, public_params dict)."""
    h = ctypes.c_void_p()
    pp = PublicParamsT()
    check(lib().inspire_setup_keyword_server(os.fsencode(path), buckets or 0, dim0 or 0, ctypes.byref(h),
                                             ctypes.byref(pp)))
    par = _public_params(pp)
    return Server._adopt(h, _plan_of(par)), par


def setup_keyword_records(keys, values, buckets=None, dim0=None):
    """KeywordServer::setup from in-memory records (lists of equal-length bytes)."""
    ks, vs = len(keys[0]), len(values[0])
    kb = ctypes.create_string_buffer(b"".join(keys), ks * len(keys))
    vb = ctypes.create_string_buffer(b"".join(values), vs * len(values))
    h = ctypes.c_void_p()
    pp = PublicParamsT()
    check(lib().inspire_setup_keyword_records(ctypes.cast(kb, ctypes.c_void_p), ks, ctypes.cast(vb, ctypes.c_void_p),
                                              vs, len(keys), buckets or 0, dim0 or 0, ctypes.byref(h),
                                              ctypes.byref(pp)))
    par = _public_params(pp)
    return Server._adopt(h, _plan_of(par)), par


def _plan_of(par):
    """Server shape from public params (the client derives it the same way,
    client/keyword.rs:60-118)."""
    kv = par["entry_size"]
    return plan_shape(par["num_items"], kv - 1, 1, buckets=par["num_items"], dim0=par["dim0"])


def _pp_struct(par):
    pp = PublicParamsT()
    pp.num_items, pp.item_size_bits, pp.dim0 = par["num_items"], par["item_size_bits"], par["dim0"]
    seed = bytes(par["cuckoo_seed"])
    if len(seed) != 16:
        raise InspireError(ERR_INVALID, "cuckoo seed must be 16 bytes")
    pp.cuckoo_seed[:] = list(seed)
    pp.num_hashes = par["num_hashes"]
    pp.key_size, pp.value_size, pp.entry_size = par["key_size"], par["value_size"], par["entry_size"]
    pp.num_records = par.get("num_records", 0)
    return pp


class KeywordClientQuery:
    """KeywordClientQuery (client/keyword.rs:26-43): the cuckoo positions and
    the serialized query."""

    def __init__(self, positions, data):
        self.positions = list(positions)
        self.data = data


class KeywordClient:
    """KeywordClient (src/pir/client/keyword.rs:45-311) in the library
    (inspire_client_*).  public_params: the dict of setup_keyword_server /
    wire.public_params_from_json.  hamming = 0 is the reference client's zero
    secret (client/base.rs:370-389); hamming > 0 a ternary secret of that
    weight.  rng_seed = 0 draws the secret randomness from the OS."""

    def __init__(self, public_params, hamming=0, rng_seed=0):
        self.public_params = dict(public_params)
        self._pp = _pp_struct(self.public_params)
        self._h = ctypes.c_void_p()
        check(lib().inspire_client_new(ctypes.byref(self._pp), hamming, rng_seed, ctypes.byref(self._h)))

    @classmethod
    def setup_from_public_params(cls, public_params, hamming=0, rng_seed=0):
        return cls(public_params, hamming, rng_seed)

    def close(self):
        if self._h:
            lib().inspire_client_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def key_size(self):
        return self.public_params["key_size"]

    def value_size(self):
        return self.public_params["value_size"]

    def secret(self):
        import numpy as np
        sk = np.zeros(2048, dtype=np.uint64)
        check(lib().inspire_client_secret(self._h, sk.ctypes.data))
        return sk

    def positions(self, key):
        key = bytes(key)
        out = (ctypes.c_int64 * self.public_params["num_hashes"])()
        kb = ctypes.create_string_buffer(key, len(key))
        check(lib().inspire_client_positions(self._h, ctypes.cast(kb, ctypes.c_void_p), len(key), out))
        return list(out)

    def query_bytes(self):
        return lib().inspire_client_query_bytes(self._h)

    def query(self, key):
        """query + serialize_query: a KeywordClientQuery (positions, bytes)."""
        key = bytes(key)
        cap = self.query_bytes()
        buf = ctypes.create_string_buffer(cap)
        pos = (ctypes.c_int64 * self.public_params["num_hashes"])()
        n = ctypes.c_size_t()
        kb = ctypes.create_string_buffer(key, len(key))
        check(lib().inspire_client_query(self._h, ctypes.cast(kb, ctypes.c_void_p), len(key), pos,
                                         ctypes.cast(buf, ctypes.c_void_p), cap, ctypes.byref(n)))
        return KeywordClientQuery(list(pos), buf.raw[:n.value])

    def serialize_query(self, query):
        return query.data

    def extract_from_serialized_response(self, key, query, response):
        """The value bytes, or None when neither a chain nor the stash holds the key."""
        key, response = bytes(key), bytes(response)
        pos = (ctypes.c_int64 * len(query.positions))(*query.positions)
        cap = max(1, self.public_params["value_size"])
        out = ctypes.create_string_buffer(cap)
        vlen, found = ctypes.c_size_t(), ctypes.c_int32()
        kb = ctypes.create_string_buffer(key, max(1, len(key)))
        rb = ctypes.create_string_buffer(response, max(1, len(response)))
        check(lib().inspire_client_extract(self._h, ctypes.cast(kb, ctypes.c_void_p), len(key), pos,
                                           ctypes.cast(rb, ctypes.c_void_p), len(response),
                                           ctypes.cast(out, ctypes.c_void_p), cap, ctypes.byref(vlen),
                                           ctypes.byref(found)))
        return out.raw[:vlen.value] if found.value else None

    extract = extract_from_serialized_response


KEM_PUBLIC_KEY_LEN = {"ml-kem-512": 800, "ml-kem-768": 1184, "ml-kem-1024": 1568}


def generate_dataset(path, kem_name, count, seed=1, threads=0):
    """DatasetGenerator::stream_to_writer to a PQKEYV2 file (generator.rs:112-169)."""
    check(lib().inspire_dataset_generate(os.fsencode(path), kem_name.encode(), count, seed, threads))


def generate_records(kem_name, count, seed=1, start=0, threads=0):
    """Records start..start+count-1 of DatasetGenerator(seed): list of (id, public key)."""
    pk_len = KEM_PUBLIC_KEY_LEN[kem_name]
    ids = ctypes.create_string_buffer(32 * count)
    keys = ctypes.create_string_buffer(pk_len * count)
    check(lib().inspire_dataset_generate_records(kem_name.encode(), seed, start, count, ctypes.cast(ids, ctypes.c_void_p),
                                                 ctypes.cast(keys, ctypes.c_void_p), threads))
    return [(ids.raw[32 * i:32 * i + 32], keys.raw[pk_len * i:pk_len * (i + 1)]) for i in range(count)]


def dataset_record_seed(seed, index):
    out = ctypes.create_string_buffer(64)
    check(lib().inspire_dataset_record_seed(seed, index, ctypes.cast(out, ctypes.c_void_p)))
    return out.raw


def mlkem_keygen(kem_name, seed64):
    """kem.rs:27-44: ML-KEM.KeyGen_internal(d, z) encapsulation key."""
    out = ctypes.create_string_buffer(KEM_PUBLIC_KEY_LEN[kem_name])
    sb = ctypes.create_string_buffer(bytes(seed64), 64)
    check(lib().inspire_mlkem_keygen(kem_name.encode(), ctypes.cast(sb, ctypes.c_void_p),
                                     ctypes.cast(out, ctypes.c_void_p), len(out)))
    return out.raw


def keccak(kind, data, out_len):
    """0 SHA3-256, 1 SHA3-512, 2 SHAKE128, 3 SHAKE256."""
    out = ctypes.create_string_buffer(max(out_len, 1))
    ib = ctypes.create_string_buffer(bytes(data), len(data) or 1)
    check(lib().inspire_keccak(kind, ctypes.cast(ib, ctypes.c_void_p), len(data), ctypes.cast(out, ctypes.c_void_p),
                               out_len))
    return out.raw[:out_len]


def cuckoo_positions(key, num_buckets):
    """CuckooHash::all_positions with the deterministic seed (cuckoo.rs:98-132)."""
    out = (ctypes.c_int64 * 2)()
    kb = ctypes.create_string_buffer(bytes(key), len(key) or 1)
    check(lib().inspire_cuckoo_positions(ctypes.cast(kb, ctypes.c_void_p), len(key), num_buckets, out))
    return [out[0], out[1]]


def siphash24(k0, k1, msg):
    mb = ctypes.create_string_buffer(bytes(msg), len(msg) or 1)
    return lib().inspire_siphash24(k0, k1, ctypes.cast(mb, ctypes.c_void_p), len(msg))


def setup_offline_device(d_db, db_rows, instances, d_bt, d_btb, d_bth, d_a_hat, stream=None):
    """Offline precomputation on caller-owned device buffers (keyword.rs:427-458)."""
    check(lib().inspire_setup_offline_device(_ptr(d_db), db_rows, instances, _ptr(d_bt), _ptr(d_btb), _ptr(d_bth),
                                             _ptr(d_a_hat), stream))


def unpack_query_device(d_query, nq, db_rows, d_y, d_z, d_rows, d_bodies, stream=None):
    check(lib().inspire_unpack_query_device(_ptr(d_query), nq, db_rows, _ptr(d_y), _ptr(d_z), _ptr(d_rows),
                                            _ptr(d_bodies), stream))


def gemv_device(d_c, d_rows, nq, d_db, db_rows, db_cols, stream=None):
    check(lib().inspire_gemv_device(_ptr(d_c), _ptr(d_rows), nq, _ptr(d_db), db_rows, db_cols, stream))


GEMV_VALU, GEMV_MFMA = 0, 1
PACKING_VALU, PACKING_MFMA = 0, 1


def gemv_mfma_scratch_bytes(db_rows):
    return lib().inspire_gemv_mfma_scratch_bytes(db_rows)


def gemv_mfma_device(d_c, d_rows, nq, d_db, db_rows, db_cols, d_scratch, stream=None):
    check(lib().inspire_gemv_mfma_device(_ptr(d_c), _ptr(d_rows), nq, _ptr(d_db), db_rows, db_cols,
                                         _ptr(d_scratch), stream))


def encode_db_device(d_out, d_in, db_rows, db_cols, interp_degree, stream=None):
    """encode_kernel.cu:263-380 on device buffers (column-major u16)."""
    check(lib().inspire_encode_db_device(_ptr(d_out), _ptr(d_in), db_rows, db_cols, interp_degree, stream))


def packing_device(d_S, d_bt, d_btb, d_bth, d_y, d_z, instances, stream=None):
    check(lib().inspire_packing_device(_ptr(d_S), _ptr(d_bt), _ptr(d_btb), _ptr(d_bth), _ptr(d_y), _ptr(d_z),
                                       instances, stream))


def post_device(d_P, d_S, count, stream=None):
    check(lib().inspire_post_device(_ptr(d_P), _ptr(d_S), count, stream))


def fold_device(d_out, d_a_hat, d_c, d_P, d_bodies, instances, d, nq, stream=None):
    check(lib().inspire_fold_device(_ptr(d_out), _ptr(d_a_hat), _ptr(d_c), _ptr(d_P), _ptr(d_bodies),
                                    instances, d, nq, stream))


def serialize_query(y_body, z_body, rows, bodies):
    """serialize_keyword_query (wire.rs:542-615), the client side of the wire
    format: a count byte, then 56-bit fields (28 bits per CRT residue) packed
    little-endian, i.e. field f at bytes [1 + 7f, 8 + 7f).

    y_body, z_body: [3][2048] condensed u64; rows: [nq][db_rows] condensed u64;
    bodies: [nq][6][2][2048] residues."""
    import numpy as np
    nq = rows.shape[0]
    m28 = np.uint64((1 << 28) - 1)

    def cond(v):
        v = np.asarray(v, dtype=np.uint64)
        return (v & m28) | ((v >> np.uint64(32)) & m28) << np.uint64(28)

    keys = np.empty((3, 2048, 2), dtype=np.uint64)
    keys[:, :, 0] = cond(y_body)
    keys[:, :, 1] = cond(z_body)
    parts = [keys.reshape(-1)]
    for h in range(nq):
        parts.append(cond(rows[h]))
        b = np.asarray(bodies[h], dtype=np.uint64)
        parts.append(((b[:, 0, :] & m28) | ((b[:, 1, :] & m28) << np.uint64(28))).reshape(-1))
    fields = np.concatenate(parts)
    raw = fields.view(np.uint8).reshape(-1, 8)[:, :7].reshape(-1)
    return bytes([nq]) + raw.tobytes()


def exported_symbols():
    """Function names declared in include/inspire_hip.h."""
    import re
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    keep = [n for n in names if n.startswith(("gpu_", "inspire_"))]
    return sorted(set(keep))
