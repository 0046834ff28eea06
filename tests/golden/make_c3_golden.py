"""Writes tests/golden/c3_full.json: per-range prime count, base-2 checksum and sum of q of
the whole batched workload (4096 ranges from bench.c3_ranges: starts uniform in [1e12, 1e18],
lengths log-uniform in [1e6, 1e9], splitmix64 seed 0x243F6A8885A308D3), computed by the CPU
oracle (whose arithmetic and prime set are pinned by tests/test_oracle.py), one range per
worker process.  About an hour on 8 cores.

usage: python tests/golden/make_c3_golden.py [PROCESSES]
"""
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def _init():
    import oracle
    oracle.scan(10 ** 18 - 100, 10 ** 18 - 1, threads=1)   # sieving primes up to 1e9, once per process


def _one(args):
    import oracle
    i, lo, hi = args
    r = oracle.scan(lo, hi, bases=[2], threads=1)
    return i, r["primes"], r["tests"], r["cks"][2], r["sumq"][2]


def main():
    from bench import c3_ranges
    nproc = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    ranges = c3_ranges()
    t = time.time()
    res = [None] * len(ranges)
    # longest ranges first, so that the pool ends evenly
    order = sorted(range(len(ranges)), key=lambda i: ranges[i][0] - ranges[i][1])
    with mp.Pool(nproc, initializer=_init) as pool:
        for n, (i, p, ts, c, q) in enumerate(pool.imap_unordered(_one, [(i,) + ranges[i] for i in order])):
            res[i] = (p, ts, c, q)
            if n % 256 == 255:
                print("%d ranges, %.0f s" % (n + 1, time.time() - t), flush=True)
    M = (1 << 64) - 1
    digest = 0
    for p, ts, c, q in res:
        digest = (digest * 0x100000001B3 ^ c ^ p) & M
    out = {"workload": "4096 ranges, starts U[1e12,1e18], lengths logU[1e6,1e9], seed 0x243F6A8885A308D3",
           "ranges": len(ranges), "primes": sum(r[0] for r in res), "tests": sum(r[1] for r in res),
           "per_range_digest": "%016x" % digest,
           "sumq2_total": "%016x" % (sum(r[3] for r in res) & M),
           "per_range_primes": [r[0] for r in res],
           "per_range_cks2": ["%016x" % r[2] for r in res],
           "oracle_seconds": round(time.time() - t, 1), "processes": nproc,
           "_source": "oracle/wf_oracle.c (pinned by tests/test_oracle.py) via tests/golden/make_c3_golden.py"}
    with open(os.path.join(HERE, "c3_full.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(out["primes"], out["per_range_digest"])


if __name__ == "__main__":
    main()
