"""Events and failure paths on the GPU, through the C ABI:

* an injected residue fault (wf_config.fault_p) must surface as the reference's FLT error
  (A001220/wieferich.c:1058-1063): an FLT event with the corrupted r and its q >= p, the pair
  counted in tests and flt and left out of the checksum and the sum of q, on every
  arithmetic path (FP64 p-adic, radix 2^28, the 32-bit limb ladders, composite bases);
* more events than the event buffer holds is an error (WF_E_OVERFLOW), never a silent
  truncation (the reference GPU tool's known failure mode, wieferich_cuda.cu:118-120);
* wf_scan_batch's events carry the index of their range, also for overlapping and repeated
  ranges (the reference's per-chunk Event list, wieferich_cuda.cu:662-675).
The expected values come from the oracle and Python's pow.
"""
import sys

import pytest

import oracle
from mathpub_amd import wf

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1
GOLD = 0x9E3779B97F4A7C15


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def corrupted(p, b):
    """The FLT event of pair (p, b) when its residue r is taken as r + 1: (q', r + 1), with
    q' = r p^-1 mod 2^64 computed from Python's pow."""
    r = pow(b, p - 1, p * p) + 1
    return ((r - 1) * pow(p, -1, 1 << 64)) & M64, r


# (arith, bases, window around the faulted prime): FP64 p-adic below 9 * 2^47 (base 2 and the
# generic window ladder), composite bases derived from prime ladders, radix 2^28 above it,
# and the 32-bit limb ladders (arith = 1) below and above 2^63
CASES = [(0, [2], 10 ** 15), (0, [3], 10 ** 15), (0, [2, 6, 10, 12], 10 ** 15), (0, [2], 3 * 10 ** 4),
         (0, [2], 10 ** 18), (0, [7], 10 ** 18), (1, [2], 10 ** 15), (1, [2], (1 << 63) + (1 << 40))]


@pytest.mark.parametrize("arith,bases,h", CASES)
def test_fault_becomes_flt_event(arith, bases, h):
    lo, hi = h, h + 200000
    # a prime whose corrupted pairs fail the FLT check (q' >= p): above 2^63 about half do not
    P = next(p for p in oracle.primes(lo + 50000, lo + 60000)[3:] if all(corrupted(p, b)[0] >= p for b in bases))
    say("fault", arith, bases, P)
    with wf.Engine(arith=arith, fault_p=P) as e:
        e.set_bases(bases)
        r = e.scan(lo, hi, near_T=1000)
        per, ev_b = e.scan_batch([(lo, hi)], near_T=1000)
    o = oracle.scan(lo, hi, bases=bases, near_T=1000)
    for x, evs in ((r, r["events"]), (per[0], [t[:5] for t in ev_b])):
        assert (x["primes"], x["tests"], x["flt"]) == (o["primes"], o["tests"], len(bases))
        flt = [t for t in evs if t[0] == wf.EV_FLT]
        assert flt == [(wf.EV_FLT, P, corrupted(P, b)[0], corrupted(P, b)[1], b) for b in bases]
        for j, b in enumerate(bases):
            q = pow(b, P - 1, P * P) // P
            assert x["cks_all"][j] == (o["cks"][b] - oracle.mix64(q ^ (P * GOLD & M64))) & M64
            assert x["sumq_all"][j] == (o["sumq"][b] - q) & M64
        others = [t for t in evs if t[0] != wf.EV_FLT]
        assert [(t[1], t[2]) for t in others] == [(t[2], t[3]) for t in o["events"]]


def test_event_overflow_is_an_error():
    """near_T above every p makes every tested prime a NEAR event: 96k of them in this range.
    With room for 65536 the scan fails with WF_E_OVERFLOW; with room for all it returns every
    one, and they equal the oracle's."""
    lo, hi = 10 ** 9, 10 ** 9 + 2 * 10 ** 6
    with wf.Engine() as e:
        with pytest.raises(wf.WFError) as ex:
            e.scan(lo, hi, near_T=1 << 62, ev_cap=1 << 16)
        assert ex.value.rc == -5
        with pytest.raises(wf.WFError) as ex:
            e.scan_batch([(lo, hi)], near_T=1 << 62, ev_cap=1 << 16)
        assert ex.value.rc == -5
        r = e.scan(lo, hi, near_T=1 << 62, ev_cap=1 << 17)
    o = oracle.scan(lo, hi, bases=[2], near_T=1 << 62, ev_cap=1 << 17)
    assert r["nev"] == o["nev"] == r["tests"] > 1 << 16
    assert [(t, p, q) for t, p, q, rr, b in r["events"]] == [(x[0], x[2], x[3]) for x in o["events"]]


def test_batch_events_carry_their_range():
    """Overlapping and repeated ranges in one batch: each range's events, selected by their
    range index, equal the events of the same range scanned alone; the list is sorted by
    (p, base, range)."""
    ranges = [(0, 4000), (1000, 5000), (0, 4000), (3500, 3600), (10 ** 15, 10 ** 15 + 3 * 10 ** 6),
              (10 ** 15 + 10 ** 6, 10 ** 15 + 2 * 10 ** 6), (10 ** 18, 10 ** 18 + 10 ** 6), (1092, 1094)]
    with wf.Engine() as e:
        e.set_bases([2, 3])
        per, ev = e.scan_batch(ranges, near_T=3000)
        alone = [e.scan(lo, hi, near_T=3000) for lo, hi in ranges]
    assert ev == sorted(ev, key=lambda t: (t[1], t[4], t[5]))
    for i, (a, r) in enumerate(zip(alone, per)):
        mine = [t[:5] for t in ev if t[5] == i]
        assert mine == a["events"], ranges[i]
        assert r["nev"] == len(mine) == a["nev"]
        assert (r["primes"], r["cks_all"], r["sumq_all"]) == (a["primes"], a["cks_all"], a["sumq_all"])
    assert [t[5] for t in ev if t[1] == 1093 and t[4] == 2] == [0, 1, 2, 7]   # FOUND in four ranges
