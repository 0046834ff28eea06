"""GPU parity at the full sizes of the single-range, batched and top-of-word workloads.

* [1e15, 1e15 + 1e11): count, checksum, sum of q, near list and solutions against the oracle's
  values for the whole range (tests/golden/c2_full.json), through wf_scan and wf_scan_batch.
* 4096 batched ranges (starts uniform in [1e12, 1e18], lengths log-uniform in [1e6, 1e9]):
  every range's prime count and base-2 checksum, and the total sum of q, against the
  oracle's values for the same ranges (tests/golden/c3_full.json, make_c3_golden.py).
* [2^64 - 2^40, 2^64): the oracle would need hours here, so two independent arithmetic
  paths must agree on every total (radix-2^28 Montgomery with the bucket sieve, against
  32-bit limbs keeping the 129th bit with 16 KiB segments and prove mode), and every listed
  strong pseudoprime is checked on the CPU: composite, and a strong probable prime to base 2
  (its BPSW count must be 0).
"""
import json
import os
import sys

import pytest

import oracle
from bench import c3_ranges
from mathpub_amd import wf

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = (1 << 64) - 1


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("arith", [0, 1], ids=["fast", "int32"])
def test_c2_full_vs_golden(arith):
    g = json.load(open(os.path.join(G, "c2_full.json")))
    say("C2 full [1e15, 1e15+1e11), arith", arith)
    with wf.Engine(arith=arith) as e:
        r = e.scan(g["lo"], g["hi"], near_T=g["near_T"])
        per, ev = e.scan_batch([(g["lo"], g["hi"])], near_T=g["near_T"])
    for x in (r, per[0]):
        assert (x["primes"], x["tests"], x["near"], x["flt"]) == (g["primes"], g["tests"], g["near"], g["flt"])
        assert "%016x" % x["cks"] == g["cks2"] and "%016x" % x["sumq"] == g["sumq2"]
    near = [[p, q if q <= p // 2 else q - p] for t, p, q, rr, b in r["events"] if t == wf.EV_NEAR]
    assert near == g["near_list"]
    assert [p for t, p, q, rr, b in r["events"] if t == wf.EV_FOUND] == g["hits"]


@pytest.mark.parametrize("arith", [0, 1], ids=["fast", "int32"])
def test_c3_full_vs_golden(arith):
    path = os.path.join(G, "c3_full.json")
    if not os.path.exists(path):
        pytest.skip("tests/golden/c3_full.json not generated")
    g = json.load(open(path))
    ranges = c3_ranges()
    assert len(ranges) == g["ranges"] == 4096
    say("C3 full, 4096 ranges, arith", arith)
    with wf.Engine(arith=arith) as e:
        per, ev = e.scan_batch(ranges)
    assert [r["primes"] for r in per] == g["per_range_primes"]
    assert ["%016x" % r["cks"] for r in per] == g["per_range_cks2"]
    assert sum(r["tests"] for r in per) == g["tests"]
    assert "%016x" % (sum(r["sumq"] for r in per) & M64) == g["sumq2_total"]
    assert all(r["flt"] == 0 for r in per)


def test_c4_full_two_arithmetics():
    lo, hi = (1 << 64) - (1 << 40), 1 << 64
    say("C4 full [2^64-2^40, 2^64), radix-2^28 with SPRP-2 list")
    with wf.Engine() as e:
        a = e.scan(lo, hi, near_T=100, psp=True)
    say("C4 full, 32-bit limbs, 16 KiB segments, prove mode")
    with wf.Engine(arith=1, seg_bytes=16384, prove=wf.PROVE_ON) as e:
        b = e.scan(lo, hi, near_T=100)
    say("primes", a["primes"], "psp", a["npsp"])
    for k in ("primes", "tests", "cks", "sumq", "near", "flt"):
        assert a[k] == b[k], k
    assert a["flt"] == 0 and a["primes"] > 2.4e10
    assert [(t, p, q) for t, p, q, r, bb in a["events"]] == [(t, p, q) for t, p, q, r, bb in b["events"]]
    assert a["npsp"] == len(a["psp"]) and a["psp"] == sorted(a["psp"])
    for n in a["psp"]:
        assert lo <= n and n & 1 and oracle.sprp2(n) and not oracle.is_prime64(n), n
    assert a["nbpsw"] == 0
