"""Writes tests/golden/c2_full.json: count, base-2 checksum, sum of q and near list of the
whole single-range workload [1e15, 1e15 + 1e11), computed by the CPU oracle (whose
arithmetic and prime set are pinned by tests/test_oracle.py).  About 4 minutes on 8 cores.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

t = time.time()
r = oracle.scan(10 ** 15, 10 ** 15 + 10 ** 11, bases=[2], near_T=100, threads=8, ev_cap=1 << 20)
out = {"lo": 10 ** 15, "hi": 10 ** 15 + 10 ** 11, "primes": r["primes"], "tests": r["tests"],
       "cks2": "%016x" % r["cks"][2], "sumq2": "%016x" % r["sumq"][2], "near_T": 100, "near": r["near"],
       "flt": r["flt"], "near_list": [[e[2], e[3] if e[3] <= e[2] // 2 else e[3] - e[2]] for e in r["events"] if e[0] == 1],
       "hits": [e[2] for e in r["events"] if e[0] == 0], "oracle_seconds_8_threads": round(time.time() - t, 1),
       "_source": "oracle/wf_oracle.c (pinned by tests/test_oracle.py) via tests/golden/make_c2_golden.py, 8 threads"}
json.dump(out, open(os.path.join(HERE, "c2_full.json"), "w"), indent=1)
print(out["primes"], out["cks2"])
