"""Writes tests/golden/respond_small.json: SHA-256 digests of the oracle's full
respond (keyword.rs:534-628) on small synthetic states.  The oracle itself is
pinned by tests/test_oracle_pins.py (reference KATs, closed forms); these
digests freeze its end-to-end output so the GPU tests on a machine without the
source tree compare against fixed vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

CASES = [
    # (db_rows, instances, d, state_seed, query_seed, nq, small)
    (2048, 2, 2, 101, 201, 2, False),
    (2048, 2, 1, 102, 202, 1, True),
    (4096, 4, 2, 103, 203, 2, False),
    (2048, 8, 8, 104, 204, 2, False),   # the 10k-key shape: instances 8, d 8, c 1
]


def main():
    out = {"generator": "tests/golden/make_golden.py", "cases": []}
    for db_rows, inst, d, ss, qs, nq, small in CASES:
        st = O.synth_state(ss, db_rows, inst)
        y, z, rows, bodies = O.synth_query(qs, db_rows, nq=nq, small=small)
        resp = O.respond_raw(st[0], db_rows, inst, d, st[1], st[2], st[3], st[4], y, z, rows, bodies)
        out["cases"].append({
            "db_rows": db_rows, "instances": inst, "d": d, "state_seed": ss, "query_seed": qs,
            "nq": nq, "small": small,
            "sha256": [hashlib.sha256(r).hexdigest() for r in resp],
            "first_words": [r[:16].hex() for r in resp],
        })
        print("case", db_rows, inst, d, "done", flush=True)
    with open(os.path.join(HERE, "respond_small.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
