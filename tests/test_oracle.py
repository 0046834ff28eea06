"""The CPU oracle pinned against the reference's own data and against mathematics.

Pins (none of them retypes the oracle's formulas):
  * wieferich_cks.h: prime counts and 17-base checksums of three windows (reference data);
  * the KNOWN solution table and pi(10^9) (wieferich.c:699-723, :1851);
  * the A306256 chunk [303630000000000, 303640000000000) base 30 (OEIS draft :21-22);
  * Python's built-in pow on random primes up to 2^64 (a library routine);
  * tests/golden/base2_windows.json, written by make_golden.py from Python integers;
  * brute-force trial division for the sieve and the SPRP-2 definition.
"""
import json
import os
import random

import pytest

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FACTS = json.load(open(os.path.join(G, "reference_facts.json")))
WINS = json.load(open(os.path.join(G, "base2_windows.json")))
M64 = (1 << 64) - 1


def test_mix64_is_splitmix_finalizer():
    # splitmix64 stream from seed 0: known first outputs of the published generator
    s = 0
    outs = []
    for _ in range(3):
        s = (s + 0x9E3779B97F4A7C15) & M64
        outs.append(oracle.mix64((s - 0x9E3779B97F4A7C15) & M64))
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_sieve_brute_force():
    def brute(lo, hi):
        return [n for n in range(lo, hi) if n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1))]
    for lo, hi in [(0, 3), (0, 1000), (2, 3), (3, 4), (1, 2), (9990, 10111), (123456, 125000)]:
        assert oracle.primes(lo, hi) == brute(lo, hi), (lo, hi)
        assert oracle.primes(lo, hi, method=1) == brute(lo, hi), (lo, hi)


def test_sieve_matches_mr_near_2_64():
    lo, hi = (1 << 64) - 200000, 1 << 64
    assert oracle.primes(lo, hi) == oracle.primes(lo, hi, method=1)
    lo = 10 ** 15
    assert oracle.primes(lo, lo + 300000) == oracle.primes(lo, lo + 300000, method=1)


def test_pi_1e9_and_base2_solutions():
    r = oracle.scan(3, 10 ** 9, bases=[2], near_T=0)
    assert r["primes"] == FACTS["pi_1e9"] - 1          # 2 is not in [3, 1e9)
    assert r["tests"] == r["primes"]
    assert [e[2] for e in r["events"] if e[0] == 0] == FACTS["known"]["2"]
    assert r["flt"] == 0


def test_known_solutions_below_1e8():
    known = {int(b): v for b, v in FACTS["known"].items()}
    bases = sorted(known)
    r = oracle.scan(0, 10 ** 8, bases=bases)
    got = sorted((e[1], e[2]) for e in r["events"] if e[0] == 0)
    want = sorted((b, p) for b in bases for p in known[b] if p < 10 ** 8)
    assert got == want
    assert r["primes"] == 5761455                       # pi(10^8)


@pytest.mark.parametrize("w", range(3))
def test_ckwin_17_bases(w):
    c = FACTS["ckwin"][w]
    r = oracle.scan(c["lo"], c["hi"], bases=FACTS["oeis_bases"])
    assert r["primes"] == c["primes"]
    assert r["flt"] == 0
    assert ["%016x" % r["cks"][b] for b in FACTS["oeis_bases"]] == c["cks"]


@pytest.mark.slow
def test_a306256_chunk():
    c = FACTS["a306256_chunk"]
    r = oracle.scan(c["lo"], c["hi"], bases=[30])
    assert r["primes"] == c["primes"]
    assert "%016x" % r["cks"][30] == c["cks"]
    assert [e[2] for e in r["events"] if e[0] == 0] == [c["solution"]]


@pytest.mark.parametrize("w", range(len(WINS)))
def test_base2_windows(w):
    g = WINS[w]
    r = oracle.scan(g["lo"], g["hi"], bases=[2], near_T=g["near_T"], psp=g["psp"] is not None)
    assert r["primes"] == g["primes"]
    assert "%016x" % r["cks"][2] == g["cks2"]
    assert "%016x" % r["sumq"][2] == g["sumq2"]
    near = [[e[2], (e[3] if e[3] <= e[2] // 2 else e[3] - e[2])] for e in r["events"] if e[0] == 1]
    assert near == g["near"]
    assert [e[2] for e in r["events"] if e[0] == 0] == g["hits"]
    if g["psp"] is not None:
        assert r["psp"] == g["psp"]


def test_quotients_vs_python_pow():
    rng = random.Random(0x243F6A8885A308D3)
    for _ in range(3000):
        bits = rng.randint(3, 64)
        p = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        while not oracle.is_prime64(p):
            p += 2
        if p >= 1 << 64:
            continue
        b = rng.choice([2, 2, 2, 3, 7, 30, 4294967291])
        if b % p == 0:
            continue
        q, r = oracle.fermat_quotient(p, b)
        R = pow(b, p - 1, p * p)
        assert r == R and q == (R - 1) // p, (p, b)


def test_edge_moduli():
    # wieferich.c:1753 edge moduli plus the ends of every limb class and of the word
    for p in [3, 5, 7, 4294967291, 4294967311, (18446744073709551557 >> 1) | 1, (1 << 63) - 25,
              (1 << 31) - 1, (1 << 61) - 1, 18446744073709551557, 18446744073709551533]:
        if not oracle.is_prime64(p):
            continue
        q, r = oracle.fermat_quotient(p, 2)
        assert r == pow(2, p - 1, p * p)


def test_montgomery_primitives():
    rng = random.Random(5)
    for _ in range(2000):
        bits = rng.randint(2, 64)
        p = (rng.getrandbits(bits) | (1 << (bits - 1)) | 1)
        if p < 3:
            continue
        N = p * p
        pinv, ninv, one = oracle.pctx(p)
        assert (p * pinv) & M64 == 1
        assert (N * ninv) & M64 == M64
        assert one == (1 << 128) % N
        x = rng.getrandbits(128) % N
        c = rng.getrandbits(32)
        assert oracle.msmall(p, x, c) == x * c % N
        Rinv = pow(1 << 128, -1, N)
        assert oracle.msqr(p, x) == x * x * Rinv % N
        assert oracle.redc(p, x) == x * Rinv % N


def test_powmod_ref_vs_pow():
    rng = random.Random(9)
    for _ in range(300):
        m = rng.getrandbits(128) | 1
        b = rng.getrandbits(32) | 2
        e = rng.getrandbits(64)
        assert oracle.powmod_ref(b, e, m) == pow(b, e, m)


def test_sprp2_definition():
    # brute force definition on all odd n < 20000, against the strong test written out
    def strong2(n):
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        x = pow(2, d, n)
        return x in (1, n - 1) or any(pow(x, 1 << r, n) == n - 1 for r in range(1, s))
    for n in range(3, 20000, 2):
        assert oracle.sprp2(n) == strong2(n), n
    # the smallest strong pseudoprimes to base 2 (OEIS A001262)
    assert [n for n in range(3, 20000, 2) if oracle.sprp2(n) and not oracle.is_prime64(n)] == \
        [2047, 3277, 4033, 4681, 8321, 15841]


def test_sumq_and_near_small_window():
    lo, hi = 3, 50000
    r = oracle.scan(lo, hi, bases=[2, 3], near_T=50)
    ps = [p for p in oracle.primes(lo, hi)]
    for b in (2, 3):
        s = 0
        near = []
        for p in ps:
            if b % p == 0:
                continue
            q = (pow(b, p - 1, p * p) - 1) // p
            s += q
            A = q if q <= p // 2 else q - p
            if q and abs(A) <= 50:
                near.append((p, A))
        assert r["sumq"][b] == s & M64
        got = [(e[2], e[3] if e[3] <= e[2] // 2 else e[3] - e[2]) for e in r["events"] if e[0] == 1 and e[1] == b]
        assert got == near


# ---- statistics for the other sieve consumers (twin pairs, runs of increasing gaps) ----

def _stats_brute(lo, hi):
    """From the definitions, over a plain list of primes (sympy): twins, L(p) by walking."""
    import sympy
    ps = list(sympy.primerange(0, hi + 3000))
    runs, first, tw, best = {}, {}, 0, (0, 0)
    for i, p in enumerate(ps):
        if p < lo or p >= hi:
            continue
        g0 = ps[i + 1] - p
        tw += g0 == 2
        if g0 > best[0]:
            best = (g0, p)
        L, gp, j = 0, 0, i
        while ps[j + 1] - ps[j] > gp:
            gp = ps[j + 1] - ps[j]
            L += 1
            j += 1
        runs[L] = runs.get(L, 0) + 1
        first.setdefault(L, p)
    return {"twins": tw, "runs": runs, "first": first, "max_gap": best[0], "max_gap_p": best[1]}


@pytest.mark.parametrize("lo,hi", [(0, 3), (0, 8), (2, 3), (3, 100), (0, 200000), (123457, 300001)])
def test_sieve_stats_brute_force(lo, hi):
    s = oracle.sieve_stats(lo, hi)
    b = _stats_brute(lo, hi)
    for k in ("twins", "runs", "first", "max_gap", "max_gap_p"):
        assert s[k] == b[k], (k, s[k], b[k])


def test_sieve_stats_reference_data():
    """[0, 10^8): the twin pairs below 10^n (A007508/README.md) for n <= 8 and the first
    primes a(1)..a(10) of A158939 (A158939/README.md), with the gaps the README lists after
    them (the last one ends the run)."""
    s = oracle.sieve_stats(0, 10 ** 8)
    assert s["primes"] == 5761455
    a7508 = FACTS["a007508"]
    for n in range(1, 9):
        assert oracle.sieve_stats(0, 10 ** n)["twins"] == a7508[n - 1], n
    for n in range(1, 11):
        a = FACTS["a158939"][str(n)]
        assert s["first"][n] == a, n
        ps = oracle.primes(a, a + 10 ** 4)
        gaps = [ps[i + 1] - ps[i] for i in range(n + 1)]
        assert gaps == FACTS["a158939_gaps"][str(n)], n
    assert max(s["runs"]) == 10
