"""GPU parity of wf_sieve_stats, the sieve exposed to the repository's other prime consumers:
twin pairs (A007508), the largest gap, and the run lengths of A158939 with their first
primes.  Against the oracle on windows at every height (several passes, prove mode on and
off), and against the reference's own data: twin counts below 10^n, a(n) of A158939 up to
a(17) and the gaps listed after a(15), a(16), a(17).
"""
import json
import os
import sys

import pytest

import oracle
from mathpub_amd import wf

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FACTS = json.load(open(os.path.join(G, "reference_facts.json")))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = wf.Engine()
    yield e
    e.close()


WINDOWS = [(0, 10 ** 7), (2, 3), (0, 100), (10 ** 12 + 11, 10 ** 12 + 10 ** 7), (10 ** 15, 10 ** 15 + 5 * 10 ** 6),
           (10 ** 18 - 3 * 10 ** 6, 10 ** 18 + 2 * 10 ** 6), (wf.STATS_MAXHI - 4 * 10 ** 6, wf.STATS_MAXHI)]


@pytest.mark.parametrize("lo,hi", WINDOWS)
def test_stats_vs_oracle(eng, lo, hi):
    say("stats", lo, hi)
    assert eng.sieve_stats(lo, hi) == oracle.sieve_stats(lo, hi)


@pytest.mark.parametrize("prove", [wf.PROVE_OFF, wf.PROVE_ON])
def test_stats_many_passes(prove):
    """Passes of ~1e6 numbers (16 KiB segments, chunk 1e6): runs and twins that cross a pass
    end are settled in the pass that owns their prime."""
    e = wf.Engine(seg_bytes=16384, chunk=10 ** 6, prove=prove)
    try:
        for lo, hi in [(10 ** 9, 10 ** 9 + 7 * 10 ** 6), (10 ** 16 + 12345, 10 ** 16 + 4 * 10 ** 6)]:
            say("stats passes", prove, lo, hi)
            assert e.sieve_stats(lo, hi) == oracle.sieve_stats(lo, hi)
    finally:
        e.close()


def test_twins_and_runs_below_1e12(eng):
    """[0, 10^12): the twin pairs below 10^n of A007508 (n <= 12), the first primes a(1)..a(13)
    of A158939 (a(13) = 968068681511), pi(10^12) = 37607912018."""
    for n in range(9, 13):
        say("stats below 10^%d" % n)
        s = eng.sieve_stats(0, 10 ** n)
        assert s["twins"] == FACTS["a007508"][n - 1], n
    assert s["primes"] == 37607912018
    for n in range(1, 14):
        assert s["first"][n] == FACTS["a158939"][str(n)], n
    assert max(s["runs"]) == 13


@pytest.mark.parametrize("n", [13, 14, 15, 16, 17])
def test_a158939_terms_in_windows(eng, n):
    """Around a(n) for n = 13..17 (up to 5.2e16): the first prime with run length n in the
    window is a(n), and the gaps after it are the reference's list."""
    a = FACTS["a158939"][str(n)]
    lo = a - 3 * 10 ** 6
    say("a158939", n, a)
    s = eng.sieve_stats(lo, a + 3 * 10 ** 6)
    assert s["first"][n] == a
    ps = eng.primes(a, a + 5000)
    gaps = [ps[i + 1] - ps[i] for i in range(n + 1)]
    assert gaps == FACTS["a158939_gaps"][str(n)]


def test_cli_stats():
    import subprocess
    wfp = os.path.join(os.path.dirname(os.path.dirname(G)), "mathpub_amd", "wf")
    p = subprocess.run([wfp, "stats", "1e6"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    o = oracle.sieve_stats(0, 10 ** 6)
    assert p.stdout.splitlines() == ["STATS [0, 1000000) primes 78498 twins 8169 max_gap 114 after 492113"] + [
        "run %2d count %d first %d" % (n, o["runs"][n], FACTS["a158939"][str(n)]) for n in range(1, 8)]
