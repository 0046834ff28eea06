"""GPU parity: the HIP path (through the C ABI) against the oracle and the golden vectors.

Bar: bit-exact prime sets, counts, checksums, sums of q, near lists and solutions.
"""
import json
import os
import random
import sys
import time

import pytest

import oracle
from mathpub_amd import wf

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WINS = json.load(open(os.path.join(G, "base2_windows.json")))
FACTS = json.load(open(os.path.join(G, "reference_facts.json")))


F64_PMAX = 9 << 47   # where the engine's FP64 p-adic class ends and the radix-2^28 class begins

def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", params=[0, 1], ids=["fast", "int32"])
def eng(request):
    e = wf.Engine(arith=request.param)
    yield e
    e.close()


def near_of(events):
    return [[p, (q if q <= p // 2 else q - p)] for t, p, q, r, b in events if t == wf.EV_NEAR]


@pytest.mark.parametrize("lo,hi", [(0, 1000), (0, 3), (2, 3), (3, 4), (7, 8), (1, 120), (29, 31), (30, 31),
                                   (999983, 1100000), (10 ** 15, 10 ** 15 + 300000),
                                   ((1 << 64) - 200000, 1 << 64), ((1 << 32) - 1000, (1 << 32) + 1000)])
def test_primes_match_oracle(eng, lo, hi):
    say("primes", lo, hi)
    assert eng.primes(lo, hi) == oracle.primes(lo, hi)


def test_quotients_vs_pow(eng):
    rng = random.Random(0x243F6A8885A308D3)
    ps = [3, 5, 7, 9, 15, 1093, 3511, (1 << 50) - 27, (1 << 50) + 55, (1 << 63) - 25, (1 << 63) + 29,
          18446744073709551557, (1 << 64) - 1, F64_PMAX - 1, F64_PMAX - 3, F64_PMAX + 1, 12 * 10 ** 14 + 7]
    for _ in range(4000):
        bits = rng.randint(2, 64)
        p = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if p >= 3:
            ps.append(p)
    say("quotients", len(ps))
    got = eng.quotients(ps)
    M64 = (1 << 64) - 1
    for p, (q, r) in zip(ps, got):
        R = pow(2, p - 1, p * p)
        assert r == R, p
        assert q == ((R - 1) * pow(p, -1, 1 << 64)) & M64, p


@pytest.mark.parametrize("w", range(len(WINS)))
def test_golden_windows(eng, w):
    g = WINS[w]
    say("golden window", g["lo"], g["hi"])
    r = eng.scan(g["lo"], g["hi"], near_T=g["near_T"], psp=g["psp"] is not None)
    assert r["primes"] == g["primes"]
    if g["psp"] is not None:
        assert r["psp"] == g["psp"] and r["npsp"] == len(g["psp"])
    assert "%016x" % r["cks"] == g["cks2"]
    assert "%016x" % r["sumq"] == g["sumq2"]
    assert r["flt"] == 0
    assert near_of(r["events"]) == g["near"]
    assert [p for t, p, q, rr, b in r["events"] if t == wf.EV_FOUND] == g["hits"]
    assert r["tests"] == r["primes"] - (1 if g["lo"] <= 2 < g["hi"] else 0)


@pytest.mark.parametrize("lo,hi", [(10 ** 15, 10 ** 15 + 2 * 10 ** 8), (10 ** 12 + 12345, 10 ** 12 + 10 ** 8 + 7),
                                   ((1 << 50) - 10 ** 7, (1 << 50) + 10 ** 7), (3, 10 ** 8),
                                   (F64_PMAX - 10 ** 7, F64_PMAX + 10 ** 7), (12 * 10 ** 14, 12 * 10 ** 14 + 10 ** 8),
                                   ((1 << 32) - 3 * 10 ** 6, (1 << 32) + 3 * 10 ** 6)])
def test_scan_vs_oracle(eng, lo, hi):
    say("scan vs oracle", lo, hi)
    r = eng.scan(lo, hi, near_T=1000)
    o = oracle.scan(lo, hi, bases=[2], near_T=1000)
    assert (r["primes"], r["tests"], r["near"], r["flt"]) == (o["primes"], o["tests"], o["near"], o["flt"])
    assert r["cks"] == o["cks"][2] and r["sumq"] == o["sumq"][2]
    assert [(p, q) for t, p, q, rr, b in r["events"]] == [(e[2], e[3]) for e in o["events"]]


def test_c1_pi_1e9(eng):
    say("C1 [3, 1e9)")
    r = eng.scan(3, 10 ** 9, near_T=100)
    assert r["primes"] == FACTS["pi_1e9"] - 1
    assert [p for t, p, q, rr, b in r["events"] if t == wf.EV_FOUND] == [1093, 3511]
    o = oracle.scan(3, 10 ** 9, bases=[2], near_T=100)
    assert r["cks"] == o["cks"][2] and r["sumq"] == o["sumq"][2] and r["near"] == o["near"]


def test_batch_equals_single(eng):
    rng = random.Random(7)
    ranges = []
    for _ in range(40):
        lo = rng.randrange(10 ** 9, 10 ** 18)
        ranges.append((lo, lo + rng.randrange(1, 3 * 10 ** 6)))
    ranges += [(0, 100), (1 << 63, (1 << 63) + 10 ** 6), ((1 << 64) - 10 ** 6, 1 << 64), ((1 << 50) - 5, (1 << 50) + 5),
               (F64_PMAX - 5, F64_PMAX + 5)]
    say("batch", len(ranges))
    res, ev = eng.scan_batch(ranges, near_T=0)
    for (lo, hi), r in zip(ranges, res):
        s = eng.scan(lo, hi)
        assert (r["primes"], r["tests"], r["cks"], r["sumq"]) == (s["primes"], s["tests"], s["cks"], s["sumq"]), (lo, hi)


@pytest.mark.parametrize("lo,hi", [(10 ** 12, 10 ** 12 + 3 * 10 ** 6), ((1 << 64) - 4 * 10 ** 6, 1 << 64),
                                   (1, 10 ** 7)])
def test_psp_vs_oracle(eng, lo, hi):
    say("psp", lo, hi)
    r = eng.scan(lo, hi, psp=True)
    o = oracle.scan(lo, hi, bases=[2], psp=True)
    assert r["psp"] == o["psp"] and r["npsp"] == o["npsp"]
    assert r["primes"] == o["primes"] and r["cks"] == o["cks"][2]


def test_autotune_keeps_results(eng):
    lo, hi = 10 ** 15, 10 ** 15 + 5 * 10 ** 8
    before = eng.scan(lo, hi)
    cfg = eng.autotune(lo, hi)
    say("autotune", cfg)
    assert cfg["seg_bytes"] in (16384, 32768, 65536, 131072)
    after = eng.scan(lo, hi)
    per, ev = eng.scan_batch([(lo, hi)])
    for x in (after, per[0]):
        assert (x["primes"], x["cks"], x["sumq"]) == (before["primes"], before["cks"], before["sumq"])


def test_autotune_batch_keeps_results():
    """wf_autotune_batch picks a configuration from a sample of the batch; every range's
    result is the same before and after."""
    rng = random.Random(71)
    ranges = []
    for _ in range(40):
        lo = rng.randrange(10 ** 12, 10 ** 18)
        ranges.append((lo, lo + int(10 ** rng.uniform(5, 8))))
    e = wf.Engine()
    try:
        before, ev0 = e.scan_batch(ranges, near_T=100)
        cfg = e.autotune_batch(ranges)
        say("autotune_batch", cfg)
        assert cfg["seg_bytes"] in (16384, 32768, 65536, 131072)
        after, ev1 = e.scan_batch(ranges, near_T=100)
        assert [(x["primes"], x["cks"], x["sumq"]) for x in before] == [(x["primes"], x["cks"], x["sumq"]) for x in after]
        assert ev0 == ev1
    finally:
        e.close()


@pytest.mark.parametrize("seg,block,bmin", [(16384, 256, 512), (65536, 1024, 262144), (131072, 1024, 524288),
                                            (131072, 512, 32768), (65536, 1024, 1 << 25)])
def test_segment_configs(seg, block, bmin):
    """Every LDS segment size, sieve workgroup and bucket threshold gives the same results.
    A threshold of 2^25 puts primes above 2^24 in the direct tier (the 64-bit wheel walk)."""
    e = wf.Engine(seg_bytes=seg, sieve_block=block, bucket_min=bmin)
    try:
        for lo, hi, psp in [(3, 2 * 10 ** 7, False), (10 ** 12 + 12345, 10 ** 12 + 2 * 10 ** 8 + 7, False),
                            ((1 << 64) - 10 ** 7, 1 << 64, True)]:
            say("segment config", seg, block, bmin, lo, hi)
            r = e.scan(lo, hi, near_T=1000, psp=psp)
            o = oracle.scan(lo, hi, bases=[2], near_T=1000, psp=psp)
            assert (r["primes"], r["tests"], r["near"]) == (o["primes"], o["tests"], o["near"])
            assert r["cks"] == o["cks"][2] and r["sumq"] == o["sumq"][2]
            if psp:
                assert r["psp"] == o["psp"]
    finally:
        e.close()


def test_serving_same_alone_and_in_mix():
    """Continuous batching: every request's result equals the same range scanned alone."""
    import time
    rng = random.Random(11)
    ranges = []
    for _ in range(48):
        lo = rng.randrange(10 ** 9, 10 ** 18)
        ranges.append((lo, lo + int(10 ** rng.uniform(5, 8.5))))
    ranges += [(0, 10 ** 6), ((1 << 50) - 10 ** 6, (1 << 50) + 10 ** 6), ((1 << 64) - 10 ** 6, 1 << 64)]
    with wf.Engine() as alone:
        want = [alone.scan(lo, hi) for lo, hi in ranges]
    with wf.Engine() as eng:
        eng.serve_start(slots=16, hbm_budget=8 << 30)
        # while it serves, the engine belongs to the server thread: other calls are refused
        for call in (lambda: eng.scan(10 ** 12, 10 ** 12 + 1000), lambda: eng.quotients([1093]),
                     lambda: eng.bpsw([7]), lambda: eng.primes(0, 100), lambda: eng.sieve_stats(0, 100),
                     lambda: eng.set_bases([2, 3])):
            with pytest.raises(wf.WFError) as ex:
                call()
            assert ex.value.rc == -6
        pending = list(enumerate(ranges))
        live, got = {}, {}
        t0 = time.time()
        while (pending or live) and time.time() - t0 < 240:
            while pending:
                i, (lo, hi) = pending[0]
                t = eng.serve_submit(lo, hi)
                if t is None:
                    break
                live[t] = i
                pending.pop(0)
            for t in list(live):
                r = eng.serve_poll(t)
                if r is not None:
                    got[live.pop(t)] = r
            time.sleep(0.002)
        eng.serve_stop()
    say("served", len(got), "of", len(ranges))
    assert len(got) == len(ranges)
    for i, w in enumerate(want):
        g = got[i]
        assert (g["primes"], g["tests"], g["cks"], g["sumq"], g["flt"], g["nev"]) == \
            (w["primes"], w["tests"], w["cks"], w["sumq"], w["flt"], w["nev"]), ranges[i]


def test_tiny_and_empty_ranges(eng):
    # one-number ranges, ranges without primes, ranges ending at 2^64, and a rejected empty range
    cases = [(2, 3), (3, 4), (4, 5), (24, 29), (90, 97), (1093, 1094), (10 ** 15 + 37, 10 ** 15 + 38),
             ((1 << 64) - 59, 1 << 64), ((1 << 64) - 1, 1 << 64), ((1 << 63) - 1, 1 << 63)]
    for lo, hi in cases:
        r = eng.scan(lo, hi)
        o = oracle.scan(lo, hi, bases=[2])
        assert (r["primes"], r["tests"], r["cks"]) == (o["primes"], o["tests"], o["cks"][2]), (lo, hi)
    with pytest.raises(wf.WFError):
        eng.scan(100, 100)
    per, ev = eng.scan_batch(cases)
    for (lo, hi), r in zip(cases, per):
        o = oracle.scan(lo, hi, bases=[2])
        assert (r["primes"], r["cks"]) == (o["primes"], o["cks"][2]), (lo, hi)


# ---- any base (the reference's -b lists) ----

OEIS = FACTS["oeis_bases"]


@pytest.fixture(scope="module", params=[0, 1], ids=["fast", "int32"])
def feng(request):
    """The fast paths (FP64 / radix-2^28 window ladders), and the 32-bit limb binary ladder,
    the double check's independent arithmetic for bases other than 2."""
    e = wf.Engine(arith=request.param)
    yield e
    e.close()


_ORACLE = {}


def oracle_scan(lo, hi, bases, near_T):
    """oracle.scan, once per argument set for both engines of feng."""
    k = (lo, hi, tuple(bases), near_T)
    if k not in _ORACLE:
        _ORACLE[k] = oracle.scan(lo, hi, bases=bases, near_T=near_T)
    return _ORACLE[k]


@pytest.mark.parametrize("w", range(3))
def test_ckwin_17_bases_reference_data(feng, w):
    """The reference's own checksum windows (wieferich_cks.h, generated by verify_wieferich.py)
    for its 17 default bases: FP64 below 9 * 2^47, radix-2^28 above."""
    c = FACTS["ckwin"][w]
    say("ckwin", c["lo"], c["hi"])
    feng.set_bases(OEIS)
    r = feng.scan(c["lo"], c["hi"])
    assert r["primes"] == c["primes"] and r["flt"] == 0
    assert ["%016x" % x for x in r["cks_all"]] == c["cks"]


def test_known_solutions_all_bases(feng):
    """Every known solution of the 23 'known' bases below 10^8, p = 2 for b = 1 (mod 4)
    included, and nothing else (OEIS values the reference's selftest checks)."""
    known = {int(b): v for b, v in FACTS["known"].items()}
    bases = sorted(known)
    feng.set_bases(bases)
    r = feng.scan(0, 10 ** 8)
    got = sorted((b, p) for t, p, q, rr, b in r["events"] if t == wf.EV_FOUND)
    want = sorted((b, p) for b in bases for p in known[b] if p < 10 ** 8)
    assert got == want
    assert r["primes"] == 5761455 and r["flt"] == 0


@pytest.mark.parametrize("lo,hi", [(3, 2 * 10 ** 6), (10 ** 12, 10 ** 12 + 2 * 10 ** 7),
                                   ((1 << 32) - 2 * 10 ** 6, (1 << 32) + 2 * 10 ** 6),
                                   ((1 << 50) - 5 * 10 ** 6, (1 << 50) + 5 * 10 ** 6),
                                   (F64_PMAX - 5 * 10 ** 6, F64_PMAX + 5 * 10 ** 6),
                                   ((1 << 63) - 5 * 10 ** 6, (1 << 63) + 5 * 10 ** 6), ((1 << 64) - 10 ** 7, 1 << 64)])
def test_any_bases_vs_oracle(feng, lo, hi):
    """Windows 5..1 (bases up to 2^32 - 1), p | b pairs, near lists and sums of q per base."""
    bases = [2, 3, 6, 10, 24, 30, 97, 1625, 1626, 65537, 4294967291]
    feng.set_bases(bases)
    say("bases vs oracle", lo, hi)
    r = feng.scan(lo, hi, near_T=100)
    o = oracle_scan(lo, hi, bases, 100)
    assert (r["primes"], r["tests"], r["near"], r["flt"]) == (o["primes"], o["tests"], o["near"], o["flt"])
    assert r["cks_all"] == [o["cks"][b] for b in bases]
    assert r["sumq_all"] == [o["sumq"][b] for b in bases]
    assert [(b, p, q) for t, p, q, rr, b in r["events"]] == [(e[1], e[2], e[3]) for e in o["events"]]


def test_quotients_any_base(feng):
    rng = random.Random(31)
    ps = [3, 5, 7, 1093, (1 << 50) - 27, (1 << 63) + 29, (1 << 64) - 59] + \
        [rng.getrandbits(rng.randint(2, 64)) | 1 for _ in range(2000)]
    ps = [p for p in ps if p >= 3]
    for b in (3, 30, 1625, 65537, 4294967291):
        got = feng.quotients(ps, base=b)
        for p, (q, r) in zip(ps, got):
            assert r == pow(b, p - 1, p * p), (b, p)


def test_batch_any_bases_equals_single(feng):
    feng.set_bases(OEIS)
    rng = random.Random(8)
    ranges = [(lo, lo + rng.randrange(1, 2 * 10 ** 6)) for lo in (rng.randrange(10 ** 9, 10 ** 18) for _ in range(12))]
    ranges += [(0, 1000), ((1 << 50) - 100, (1 << 50) + 100)]
    per, ev = feng.scan_batch(ranges)
    for (lo, hi), r in zip(ranges, per):
        s = feng.scan(lo, hi)
        assert (r["primes"], r["tests"], r["cks_all"], r["sumq_all"]) == (s["primes"], s["tests"], s["cks_all"],
                                                                         s["sumq_all"]), (lo, hi)


@pytest.mark.parametrize("bases", [OEIS, [2, 3, 4, 6, 8, 9, 12, 16, 18, 24, 27, 32],
                                   [6, 10, 14, 15, 21, 35, 1625, 65535, 4294967295],
                                   [3, 5, 15, 30, 1626, 65537 * 3, 4294967291]])
def test_composite_bases_derived_equal_direct(bases):
    """Composite bases from their prime factors' quotients (the default whenever it needs
    fewer ladders) against one ladder per base: bit-exact totals, sums and events, from p = 3
    (p dividing ladder primes and bases) across 2^32 and 2^50 to the FP64 class end at
    9 * 2^47, single and batched."""
    auto, off = wf.Engine(), wf.Engine(derive=wf.DERIVE_OFF)
    try:
        for e in (auto, off):
            e.set_bases(bases)
        for lo, hi in [(3, 3 * 10 ** 6), (10 ** 12 + 1, 10 ** 12 + 3 * 10 ** 6),
                       ((1 << 32) - 10 ** 6, (1 << 32) + 10 ** 6), ((1 << 50) - 2 * 10 ** 6, (1 << 50) + 10 ** 6),
                       (F64_PMAX - 2 * 10 ** 6, F64_PMAX + 10 ** 6)]:
            say("derived vs direct", bases[:4], lo, hi)
            a = auto.scan(lo, hi, near_T=1000)
            b = off.scan(lo, hi, near_T=1000)
            assert (a["primes"], a["tests"], a["near"], a["flt"], a["cks_all"], a["sumq_all"], a["nev"]) == \
                (b["primes"], b["tests"], b["near"], b["flt"], b["cks_all"], b["sumq_all"], b["nev"]), (lo, hi)
            assert a["events"] == b["events"]
        rng = random.Random(9)
        ranges = [(lo, lo + rng.randrange(1, 10 ** 6)) for lo in (rng.randrange(10 ** 6, F64_PMAX) for _ in range(40))]
        pa, ea = auto.scan_batch(ranges, near_T=1000)
        pb, eb = off.scan_batch(ranges, near_T=1000)
        assert [(x["tests"], x["cks_all"], x["sumq_all"]) for x in pa] == [(x["tests"], x["cks_all"], x["sumq_all"])
                                                                           for x in pb]
        assert ea == eb
    finally:
        auto.close()
        off.close()


def test_int32_any_base_batch_and_serving():
    """The 32-bit limb ladders for bases other than 2 through wf_scan_batch and serving, against
    the fast paths: every class (2, 3 and 4 limbs, and the 4 limbs with a carry bit above 2^63)."""
    bases = [3, 5, 12, 4294967291]
    rng = random.Random(41)
    ranges = [(3, 10 ** 5), ((1 << 31) - 10 ** 5, (1 << 31) + 10 ** 5), ((1 << 47) - 10 ** 5, (1 << 47) + 10 ** 5),
              ((1 << 63) - 10 ** 5, (1 << 63) + 10 ** 5), ((1 << 64) - 10 ** 5, 1 << 64)]
    ranges += [(lo, lo + rng.randrange(1, 10 ** 5)) for lo in (rng.randrange(1, 1 << 64) for _ in range(20))]
    ranges = [(lo, hi if hi < 1 << 64 else 0) for lo, hi in ranges]
    with wf.Engine() as fast, wf.Engine(arith=1) as m32:
        for e in (fast, m32):
            e.set_bases(bases)
        pa, ea = fast.scan_batch(ranges, near_T=10 ** 4)
        pb, eb = m32.scan_batch(ranges, near_T=10 ** 4)
        m32.serve_start(slots=8, hbm_budget=2 << 30)
        tickets = [m32.serve_submit(lo, hi or 1 << 64, near_T=10 ** 4) for lo, hi in ranges[:6]]
        served = {}
        t0 = time.time()
        while len(served) < len(tickets) and time.time() - t0 < 60:
            for t in tickets:
                if t not in served:
                    r = m32.serve_poll(t)
                    if r is not None:
                        served[t] = r
            time.sleep(0.002)
        m32.serve_stop()
    keys = ("primes", "tests", "near", "flt", "cks_all", "sumq_all", "nev")
    assert [[x[k] for k in keys] for x in pa] == [[x[k] for k in keys] for x in pb]
    assert ea == eb
    for i, t in enumerate(tickets):
        assert [served[t][k] for k in keys] == [pa[i][k] for k in keys], ranges[i]
