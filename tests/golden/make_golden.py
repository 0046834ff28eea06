"""Writes tests/golden/base2_windows.json: base-2 golden vectors from Python integers.

Independent of the oracle and of the engine: primes come from a small-prime sieve of
each window plus deterministic Miller-Rabin (12 prime bases <= 37, exact below
3.18e23), quotients from Python's built-in pow(2, p - 1, p * p), strong probable primes
from pow as well.  The checksum definition is the one of A001220/wieferich.c:133-143
(mix64 of q ^ p * 0x9E3779B97F4A7C15, summed mod 2^64).

Run: python3 tests/golden/make_golden.py   (about a minute)
"""
import json
import os

M64 = (1 << 64) - 1
GOLD = 0x9E3779B97F4A7C15
MR = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]


def mix64(z):
    z = (z + GOLD) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def small_primes(n):
    s = bytearray([1]) * n
    s[0:2] = b"\0\0"
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = bytearray(len(range(i * i, n, i)))
    return [i for i in range(n) if s[i]]


SMALL = small_primes(1 << 16)


def strong(n, a):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, n)
    if x == 1 or x == n - 1:
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def is_prime(n):
    if n < 2:
        return False
    for q in MR:
        if n % q == 0:
            return n == q
    return all(strong(n, a) for a in MR)


def window(lo, hi, near_T=100, want_psp=False):
    """primes, base-2 cks, sum q, near list [(p, A)], hits, psp list for [lo, hi)."""
    n = hi - lo
    alive = bytearray([1]) * n
    for q in SMALL:
        if q * q >= hi:
            break
        first = max(q * q, (lo + q - 1) // q * q)
        if first < hi:
            alive[first - lo::q] = bytearray(len(range(first - lo, n, q)))
    count = cks = sumq = 0
    near, hits, psp = [], [], []
    for i in range(n):
        x = lo + i
        if x < 2:
            continue
        prime = alive[i] and (x < (1 << 32) or is_prime(x))
        if not prime:
            if want_psp and x % 2 == 1 and x > 1 and strong(x, 2):
                psp.append(x)
            continue
        count += 1
        if x == 2:
            continue
        r = pow(2, x - 1, x * x)
        assert (r - 1) % x == 0
        q = (r - 1) // x
        cks = (cks + mix64(q ^ ((x * GOLD) & M64))) & M64
        sumq = (sumq + q) & M64
        if q == 0:
            hits.append(x)
        else:
            A = q if q <= x // 2 else q - x
            if abs(A) <= near_T:
                near.append([x, A])
    return {"lo": lo, "hi": hi, "primes": count, "cks2": "%016x" % cks, "sumq2": "%016x" % sumq,
            "near_T": near_T, "near": near, "hits": hits, "psp": psp if want_psp else None}


WINDOWS = [
    (3, 2000000, False),                            # small p: 2-limb Montgomery, hits 1093, 3511
    (1, 1000000, True),                             # strong pseudoprimes to base 2 below 10^6
    (10 ** 15, 10 ** 15 + 4000000, False),          # the first checksum window of wieferich_cks.h
    (10 ** 18, 10 ** 18 + 2000000, False),          # the second
    (2 ** 63 - 2000000, 2 ** 63, False),            # the third
    (2 ** 48 - 1000000, 2 ** 48 + 1000000, False),  # N crosses 2^96
    (2 ** 64 - 1000000, 2 ** 64, True),             # top of the word (beyond the reference)
]


def main():
    out = [window(lo, hi, want_psp=psp) for lo, hi, psp in WINDOWS]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "base2_windows.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    for w in out:
        print(w["lo"], w["hi"], w["primes"], w["cks2"], len(w["near"]), w["hits"], None if w["psp"] is None else len(w["psp"]))


if __name__ == "__main__":
    main()
