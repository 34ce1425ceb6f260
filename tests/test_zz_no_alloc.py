"""inspire_respond_device_noalloc (the respond that allocates nothing: its
working buffers live in the library's per-GPU scratch array) against
inspire_respond_device, through the C ABI.  Calls run back to back on one
stream: a 1M-key query (352 outputs, 2 hash queries), then a smaller one (8
outputs, 1 hash query), then the large one again, each byte-identical to the
slot-based respond of the same query; and a request beyond the array is
refused before any device work."""
import ctypes
import sys

import numpy as np
import pytest

import inspire_amd as IA
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _stage(torch, q):
    return torch.from_numpy(np.frombuffer(q, dtype=np.uint8).copy()).cuda()


def test_noalloc_respond_equals_respond_device():
    import torch
    assert torch.cuda.is_available()
    L = IA.lib()
    big = IA.Server(2048, 352, 32, lean=True)
    big.fill_synthetic(11)
    small = IA.Server(2048, 8, 8)
    small.fill_synthetic(3)
    cases = [("1M keys, 352 outputs, nq 2", big, O.synth_query(71, 2048, nq=2), 2),
             ("8 outputs, nq 1", small, O.synth_query(72, 2048, nq=1), 1),
             ("1M keys again", big, O.synth_query(73, 2048, nq=2, small=True), 2)]
    assert IA.noalloc_scratch_bytes(2048, 352, 32, 2) <= IA.NOALLOC_SCRATCH_BYTES
    st = torch.cuda.Stream()
    for name, srv, parts, nq in cases:
        say("noalloc:", name)
        q = O.serialize_query(*parts)
        dq = _stage(torch, q)
        blob = nq * srv.c * O.RESP_BYTES
        ref = torch.zeros(blob, dtype=torch.uint8, device="cuda")
        got = torch.full((blob,), 0xA5, dtype=torch.uint8, device="cuda")
        srv.respond_device(dq, len(q), nq, ref, st.cuda_stream)
        rc = L.inspire_respond_device_noalloc(srv.h, ctypes.c_void_p(dq.data_ptr()), len(q), nq,
                                              ctypes.c_void_p(got.data_ptr()), ctypes.c_void_p(st.cuda_stream))
        assert rc == IA.OK, IA.lib().inspire_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), name
    big.close()
    small.close()
    torch.cuda.empty_cache()


def test_noalloc_refuses_requests_beyond_the_array():
    import torch
    rows, inst, d, nq = 2048, 64, 32, 100
    need = IA.noalloc_scratch_bytes(rows, inst, d, nq, max_queries=nq)
    assert need > IA.NOALLOC_SCRATCH_BYTES
    say("noalloc: %d hash queries need %d bytes" % (nq, need))
    srv = IA.Server(rows, inst, d, max_queries=nq)
    srv.fill_synthetic(5)
    qlen = 1 + 7 * (3 * O.N * 2 + nq * (rows + 6 * O.N))
    dq = torch.zeros(qlen, dtype=torch.uint8, device="cuda")
    out = torch.zeros(nq * srv.c * O.RESP_BYTES, dtype=torch.uint8, device="cuda")
    with pytest.raises(IA.InspireError) as e:
        srv.respond_device_noalloc(dq, qlen, nq, out)
    assert e.value.code == IA.ERR_INVALID and "scratch" in str(e.value)
    srv.close()
    torch.cuda.empty_cache()
