"""Generate a static ML-KEM dataset for keyword PIR (src/bin/gen_dataset.rs):
python -m inspire_amd.gen_dataset --kem ml-kem-768 --count N --out FILE [--seed S]
[--threads T] [--buckets B] [--dim0 D] [--list [N]] [--list-keys [N]].

The records come from the library's DatasetGenerator (inspire_dataset_generate);
before generating, the server shape and setup memory estimate for the file are
printed (KeywordServer::check_setup_memory_for_shape, keyword.rs:298-364)."""
import argparse
import os
import sys

from . import KEM_PUBLIC_KEY_LEN, generate_dataset, plan_shape

ALL = 18446744073709551615


def _gib(n):
    return "%.2f GiB" % (n / (1024.0 ** 3))


def _records(path):
    with open(path, "rb") as f:
        head = f.read(9)
        f.read(head[8])
        count = int.from_bytes(f.read(4), "little")
        pk_len = int.from_bytes(f.read(2), "little")
        for _ in range(count):
            rec = f.read(32 + pk_len)
            yield rec[:32], rec[32:]


def main(argv=None):
    ap = argparse.ArgumentParser(description="Generate a static ML-KEM dataset for keyword PIR")
    ap.add_argument("--kem", default="ml-kem-768", choices=sorted(KEM_PUBLIC_KEY_LEN))
    ap.add_argument("--count", type=int, required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--batch-size", type=int, default=None, help="accepted for compatibility (batches of 100,000)")
    ap.add_argument("--threads", type=int, default=0, help="generation threads (default: all hardware threads)")
    ap.add_argument("--buckets", type=int, default=None)
    ap.add_argument("--dim0", type=int, default=None)
    ap.add_argument("--list", type=int, nargs="?", const=ALL, default=None)
    ap.add_argument("--list-keys", type=int, nargs="?", const=ALL, default=None)
    a = ap.parse_args(argv)
    pk_len = KEM_PUBLIC_KEY_LEN[a.kem]
    plan = plan_shape(a.count, 32, pk_len, buckets=a.buckets, dim0=a.dim0)
    print("server shape: buckets %d, dim0 %d, db_rows %d, db_cols %d, instances %d" % (
        plan["num_items"], plan["dim0"], plan["db_rows"], plan["db_cols"], plan["instances"]))
    print("estimated GPU setup memory: %d bytes (%s)" % (plan["peak_bytes"], _gib(plan["peak_bytes"])))
    generate_dataset(a.out, a.kem, a.count, seed=a.seed, threads=a.threads)
    print("wrote %d %s records (%d bytes) to %s" % (a.count, a.kem, os.path.getsize(a.out), a.out))
    for n, keys in ((a.list, False), (a.list_keys, True)):
        if n is None:
            continue
        for i, (rid, pk) in enumerate(_records(a.out)):
            if i >= n:
                break
            print(rid.hex() + (" " + pk.hex() if keys else ""))
    return 0


if __name__ == "__main__":
    sys.exit(main())
