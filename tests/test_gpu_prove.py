"""GPU parity of the BPSW engine and of prove mode (partial sieve + proof).

* wf_bpsw against sympy (isprime, is_strong_lucas_prp) and the strong-test definition;
* prove mode ON against OFF (and against the oracle): bit-exact primes, tests, checksums,
  sums of q, near lists and events, single and batched, several bases, bucket thresholds
  and segment sizes;
* the strong pseudoprimes of the top of the word all fail the Lucas test (no BPSW
  pseudoprime below 2^64).
"""
import random
import sys

import pytest

import oracle
from mathpub_amd import wf

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def strong2(n):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


@pytest.fixture(scope="module")
def engines():
    on = wf.Engine(prove=wf.PROVE_ON)
    off = wf.Engine(prove=wf.PROVE_OFF)
    yield on, off
    on.close()
    off.close()


def test_bpsw_flags_vs_sympy(engines):
    import sympy
    from sympy.ntheory.primetest import is_strong_lucas_prp
    on, _ = engines
    rng = random.Random(5)
    ns = list(range(0, 5000)) + [2047, 3277, 4033, 4681, 8321, 5459, 5777, 10877, 16109, 18971,
                                 3825123056546413051, 18446744066047760377, (1 << 64) - 1, (1 << 64) - 59,
                                 4294967291 ** 2, 4294967279 * 4294967291]
    ns += [rng.getrandbits(64) for _ in range(20000)] + [rng.getrandbits(64) | 1 for _ in range(20000)]
    ns += [q for q in (rng.randrange(1 << 62, 1 << 64) | 1 for _ in range(30000)) if sympy.isprime(q)][:2000]
    say("bpsw", len(ns))
    fl = on.bpsw(ns)
    for n, f in zip(ns, fl):
        if n < 3 or n % 2 == 0:
            assert f == (4 if n == 2 else 0), n
            continue
        assert bool(f & wf.BPSW_PRIME) == sympy.isprime(n), n
        assert bool(f & wf.BPSW_SPRP2) == strong2(n), n
        assert bool(f & wf.BPSW_LUCAS) == is_strong_lucas_prp(n), n


def _key(r):
    return (r["primes"], r["tests"], r["near"], r["flt"], r["cks_all"], r["sumq_all"], r["nev"])


RANGES = [(10 ** 11, 10 ** 11 + 3 * 10 ** 6), (10 ** 15 + 7, 10 ** 15 + 4 * 10 ** 6 + 11),
          ((1 << 50) - 10 ** 6, (1 << 50) + 10 ** 6), (10 ** 18 + 1, 10 ** 18 + 2 * 10 ** 6),
          ((1 << 64) - 3 * 10 ** 6, 1 << 64), (123456789012345678, 123456789012345678 + 987654)]


@pytest.mark.parametrize("lo,hi", RANGES)
def test_prove_on_equals_off(engines, lo, hi):
    on, off = engines
    say("prove on/off", lo, hi)
    a = on.scan(lo, hi, near_T=5000)
    assert on.last_timing()["proved"] > 0 or a["primes"] == 0
    b = off.scan(lo, hi, near_T=5000)
    assert off.last_timing()["proved"] == 0
    assert _key(a) == _key(b)
    assert a["events"] == b["events"]


@pytest.mark.parametrize("lo,hi", [(10 ** 15 + 7, 10 ** 15 + 2 * 10 ** 6 + 11), ((1 << 64) - 10 ** 6, 1 << 64)])
def test_prove_vs_oracle(engines, lo, hi):
    on, _ = engines
    say("prove vs oracle", lo, hi)
    r = on.scan(lo, hi, near_T=1000)
    o = oracle.scan(lo, hi, bases=[2], near_T=1000)
    assert (r["primes"], r["tests"], r["near"]) == (o["primes"], o["tests"], o["near"])
    assert r["cks"] == o["cks"][2] and r["sumq"] == o["sumq"][2]


def test_prove_primes_lists(engines):
    on, off = engines
    for lo, hi in [(10 ** 13, 10 ** 13 + 10 ** 6), ((1 << 64) - 10 ** 6, 1 << 64), (10 ** 17, 10 ** 17 + 5 * 10 ** 5)]:
        say("prove primes", lo, hi)
        assert on.primes(lo, hi) == off.primes(lo, hi)


def test_prove_batched_mixed_heights(engines):
    """Many short ranges at every height in one batched call: prove and sieved pieces side by
    side (AUTO picks per piece), equal to the fully sieved results."""
    on, off = engines
    rng = random.Random(23)
    ranges = []
    for _ in range(200):
        lo = rng.randrange(10 ** 9, (1 << 64) - 10 ** 9)
        ranges.append((lo, lo + int(10 ** rng.uniform(3, 6.5))))
    ranges += [(10 ** 15, 10 ** 15 + 3 * 10 ** 8)]          # long: sieved in AUTO
    auto = wf.Engine()
    try:
        say("prove batched", len(ranges))
        pa, ea = auto.scan_batch(ranges, near_T=100)
        assert auto.last_timing()["proved"] > 0
        po, eo = on.scan_batch(ranges, near_T=100)
        pf, ef = off.scan_batch(ranges, near_T=100)
        assert [_key(x) for x in pa] == [_key(x) for x in pf]
        assert [_key(x) for x in po] == [_key(x) for x in pf]
        assert ea == ef and eo == ef
    finally:
        auto.close()


@pytest.mark.parametrize("seg,bmin", [(16384, 512), (131072, 32768), (65536, 1048576)])
def test_prove_configs_and_bases(seg, bmin):
    """Bucket thresholds from 512 (a tenth of the numbers survive) to 2^20, segment sizes 16 to
    128 KiB, several bases: prove ON equals OFF."""
    on = wf.Engine(seg_bytes=seg, bucket_min=bmin, prove=wf.PROVE_ON)
    off = wf.Engine(seg_bytes=seg, bucket_min=bmin, prove=wf.PROVE_OFF)
    try:
        for e in (on, off):
            e.set_bases([2, 3, 5, 30, 1625])
        for lo, hi in [(10 ** 14 + 3, 10 ** 14 + 2 * 10 ** 6), (10 ** 18, 10 ** 18 + 10 ** 6)]:
            say("prove config", seg, bmin, lo, hi)
            a = on.scan(lo, hi, near_T=100)
            b = off.scan(lo, hi, near_T=100)
            assert _key(a) == _key(b) and a["events"] == b["events"]
    finally:
        on.close()
        off.close()


def test_top_of_word_psp_fail_lucas(engines):
    """The base-2 strong pseudoprimes below 2^64 found by the sieve all fail the strong Lucas
    test: BPSW has no pseudoprime there (the psp pass always sieves fully)."""
    on, _ = engines
    lo = (1 << 64) - 4 * 10 ** 6
    r = on.scan(lo, 1 << 64, psp=True)
    o = oracle.scan(lo, 1 << 64, bases=[2], psp=True)
    assert r["psp"] == o["psp"] and r["npsp"] == o["npsp"]
    assert r["nbpsw"] == 0
