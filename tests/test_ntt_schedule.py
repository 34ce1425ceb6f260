"""The reduction schedule of the device's forward NTT passes, on the CPU.

The gfx950 forward transforms (ntt_device.hpp, fwd_stage_reduction) skip the
reduction of x in the Cooley-Tukey butterflies of stages 1-7, take x below 2q
with three conditional subtractions (8q, 4q, 2q) in stage 8 and run Harvey's
butterflies after that.  This restates that schedule in exact integer
arithmetic and checks, for inputs below 2q, that no value reaches 2^32 (the
device works in 32-bit registers) and that the result is the oracle's
forward NTT (ntt.rs:211-357).  The GPU parity tests check the device code
itself; this pins the bound the schedule relies on, at its worst-case inputs.
"""
import numpy as np
import pytest

import oracle as O

N = O.N
TWO32 = 1 << 32


def lazy_forward(a, c, tables):
    """The device schedule for one polynomial over prime c; returns (the
    transform after the device's final two conditional subtractions, largest
    intermediate value)."""
    q = (O.Q0, O.Q1)[c]
    fw = tables[c, 0].astype(np.uint64)
    fwp = tables[c, 1].astype(np.uint64)
    v = np.array(a, dtype=np.uint64)
    peak = int(v.max())
    m, stage = 1, 0
    while m < N:
        stage += 1
        t = N // (2 * m)
        blk = v.reshape(m, 2, t)
        x, y = blk[:, 0, :].copy(), blk[:, 1, :].copy()
        w, wp = fw[m:2 * m, None], fwp[m:2 * m, None]
        subs = () if stage <= 7 else (8, 4, 2) if stage == 8 else (2,)
        for k in subs:
            x = np.where(x >= k * q, x - k * q, x)
        qt = (y * wp) >> np.uint64(32)            # mulhi32(y, w')
        prod = w * y - qt * np.uint64(q)          # Shoup product, exact
        assert int(prod.max()) < 2 * q
        xo, yo = x + prod, x + np.uint64(2 * q) - prod
        peak = max(peak, int(xo.max()), int(yo.max()))
        assert peak < TWO32, f"stage {stage} reaches {peak}"
        blk[:, 0, :], blk[:, 1, :] = xo, yo
        m *= 2
    # the device's final reduction assumes values below 4q
    for k in (2, 1):
        v = np.where(v >= k * q, v - k * q, v)
    return v, peak


def test_bound_constants():
    # stage 7 ends below 2q + 14q = 16q; both primes keep that under 2^32
    assert 16 * O.Q0 < TWO32 and 16 * O.Q1 < TWO32
    # any x < 16q is below 2q after subtracting 8q, 4q and 2q conditionally
    for q in (O.Q0, O.Q1):
        for x in (0, 2 * q - 1, 2 * q, 8 * q - 1, 8 * q, 16 * q - 1):
            for k in (8, 4, 2):
                x = x - k * q if x >= k * q else x
            assert x < 2 * q


@pytest.mark.parametrize("kind", ["max_2q", "max_q", "random_2q", "digits"])
def test_lazy_forward_matches_oracle(kind):
    tables = O.ntt_tables()
    rng = np.random.default_rng(11)
    polys = np.zeros((2, N), dtype=np.uint64)
    lazy_in = np.zeros((2, N), dtype=np.uint64)
    for c, q in enumerate((O.Q0, O.Q1)):
        if kind == "max_2q":
            x = np.full(N, 2 * q - 1, dtype=np.uint64)
        elif kind == "max_q":
            x = np.full(N, q - 1, dtype=np.uint64)
        elif kind == "random_2q":
            x = rng.integers(0, 2 * q, N, dtype=np.uint64)
        else:  # the fold's gadget digits, below 2^19
            x = rng.integers(0, 1 << 19, N, dtype=np.uint64)
        lazy_in[c] = x
        polys[c] = x % np.uint64(q)
    want = O.ntt_forward(polys[None])[0]
    for c in range(2):
        got, peak = lazy_forward(lazy_in[c], c, tables)
        assert np.array_equal(got, want[c])
        assert peak < TWO32
