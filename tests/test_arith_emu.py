"""The kernel arithmetic of wf_arith.h, compiled for the host (IEEE doubles with FMA
behave as the GPU's), against Python's pow: FP64 p-adic path and every Montgomery width,
including exponents padded with leading zeros as in a wavefront of mixed lengths."""
import ctypes
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "emu", "_build", "arith_emu.so")


@pytest.fixture(scope="module")
def emu():
    os.makedirs(os.path.dirname(SO), exist_ok=True)
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma", "-shared", "-fPIC",
                           "-I" + os.path.join(ROOT, "mathpub_amd", "csrc"), "-o", SO,
                           os.path.join(HERE, "emu", "arith_emu.cpp")])
    L = ctypes.CDLL(SO)
    P = ctypes.POINTER(ctypes.c_uint64)
    L.emu_fp64.argtypes = [ctypes.c_uint64, P, P]
    L.emu_mont.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, P, P]
    L.emu_sprp2.argtypes = [ctypes.c_uint64]
    L.emu_psp_possible.argtypes = [ctypes.c_uint64]
    L.emu_m28.argtypes = [ctypes.c_uint64, ctypes.c_int, P, P]
    L.emu_fp64x2.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]
    L.emu_fp64c.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]
    L.emu_m28c.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, P, P]
    return L


def call(fn, *args):
    hi, lo = ctypes.c_uint64(), ctypes.c_uint64()
    fn(*args, ctypes.byref(hi), ctypes.byref(lo))
    return (hi.value << 64) | lo.value


def samples(lo_bits, hi_bits, n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        bits = rng.randint(lo_bits, hi_bits)
        p = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        out.append(p)
    return out


def test_fp64_padic(emu):
    ps = [3, 5, 7, 9, 1093, 3511, (1 << 50) - 1, (1 << 50) - 27, (1 << 49) + 1, 10 ** 15 + 37] + \
        samples(2, 50, 20000, 1)
    for p in ps:
        if p < 3:
            continue
        assert call(emu.emu_fp64, p) == pow(2, p - 1, p * p), p


def test_fp64_two_ladders(emu):
    """The kernel's two-ladder form: pairs of primes of different sizes, with the
    wavefront's exponent length (the largest of the pair, or more), small p whose leading
    window digit 2^d exceeds p, and p at the top of the class."""
    ps = [3, 5, 7, 11, 13, 31, 37, 1093, 3511, (1 << 50) - 27, (1 << 49) + 1] + samples(2, 50, 6000, 7)
    rng = random.Random(11)
    out = (ctypes.c_uint64 * 4)()
    for p0 in ps:
        if p0 < 3:
            continue
        p1 = rng.choice(ps)
        if p1 < 3:
            p1 = 3
        eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
        for e in (eb, min(50, eb + rng.randint(0, 7))):
            emu.emu_fp64x2(p0, p1, e, out)
            assert (out[0] << 64 | out[1]) == pow(2, p0 - 1, p0 * p0), (p0, p1, e)
            assert (out[2] << 64 | out[3]) == pow(2, p1 - 1, p1 * p1), (p0, p1, e)


@pytest.mark.parametrize("L,full,lo,hi", [(2, 0, 2, 31), (3, 0, 32, 47), (4, 0, 48, 63), (4, 1, 64, 64),
                                          (4, 1, 63, 64)])
def test_montgomery(emu, L, full, lo, hi):
    ps = samples(lo, hi, 5000, L * 10 + full)
    if L == 4 and full:
        ps += [(1 << 64) - 1, (1 << 64) - 59, (1 << 63) + 1]
    if L == 2:
        ps += [3, 5, 7, 9, (1 << 31) - 1]
    for p in ps:
        if p < 3:
            continue
        eb = (p - 1).bit_length()
        assert call(emu.emu_mont, L, full, p, eb) == pow(2, p - 1, p * p), (L, p)
        # a longer wave-wide exponent length only adds skipped leading bits
        assert call(emu.emu_mont, L, full, p, min(64, eb + 3)) == pow(2, p - 1, p * p), (L, p)


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


def test_sprp2_64(emu):
    ns = list(range(3, 20001, 2)) + [2047, 3277, 4033, 4681, 8321, (1 << 64) - 59, (1 << 64) - 1, (1 << 63) + 1,
                                     3825123056546413051, 318665857834031151167461 % (1 << 64) | 1]
    rng = random.Random(3)
    ns += [rng.getrandbits(64) | 1 for _ in range(3000)] + [rng.getrandbits(40) | 1 for _ in range(1000)]
    for n in ns:
        if n < 3:
            continue
        assert bool(emu.emu_sprp2(n)) == strong2(n), n


def test_radix28_montgomery(emu):
    ps = samples(47, 64, 8000, 99) + [(1 << 47) + 1, (1 << 50) - 27, (1 << 63) - 25, (1 << 63) + 29,
                                      (1 << 64) - 59, (1 << 64) - 1, (1 << 64) - 3, 10 ** 18 + 9]
    for p in ps:
        if p < (1 << 47):
            continue
        eb = (p - 1).bit_length()
        assert call(emu.emu_m28, p, eb) == pow(2, p - 1, p * p), p
        assert call(emu.emu_m28, p, min(64, eb + 2)) == pow(2, p - 1, p * p), p


def test_psp_filter_keeps_every_strong_pseudoprime(emu):
    """psp_possible is a necessary condition: it holds for every base-2 strong pseudoprime
    below 2e5 (brute force), for the one found below 2^64 in the top-of-word slice, and
    whenever the exact test passes on random odd numbers; it rejects a good share of
    ordinary composites (else it would be pointless)."""
    small = [n for n in range(9, 200000, 2) if not _isprime(n) and strong2(n)]
    assert small[:4] == [2047, 3277, 4033, 4681]
    for n in small + [18446744066047760377]:
        assert emu.emu_psp_possible(n), n
    rng = random.Random(5)
    rejected = comps = 0
    for _ in range(20000):
        n = rng.randrange(1 << 63, 1 << 64) | 1
        if _isprime(n):
            continue
        comps += 1
        ok = emu.emu_psp_possible(n)
        if emu.emu_sprp2(n):
            assert ok, n
        rejected += not ok
    assert 0.4 < rejected / comps < 0.8


def _isprime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    return all(_mr(n, a) for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37))


def _mr(n, a):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


BASES = [2, 3, 5, 6, 7, 10, 12, 13, 14, 15, 17, 18, 19, 20, 22, 23, 26, 30, 31, 1625, 1626, 65535, 65536,
         (1 << 31) - 1, (1 << 32) - 1]


def test_fp64_any_base(emu):
    """Any base 2 <= b < 2^32 through the FP64 two-ladder form (window 5..1 by base), p < 2^50,
    including p | b pairs (r = 0) and small p below the multiplier."""
    rng = random.Random(21)
    ps = [3, 5, 7, 11, 13, 1093, 3511, (1 << 50) - 27] + samples(2, 50, 1500, 8)
    out = (ctypes.c_uint64 * 4)()
    for b in BASES:
        for p0 in ps:
            if p0 < 3:
                continue
            p1 = rng.choice(ps) | 1
            if p1 < 3:
                p1 = 3
            eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
            eb = min(50, eb + rng.randint(0, 3))
            emu.emu_fp64c(b, p0, p1, eb, out)
            assert (out[0] << 64 | out[1]) == pow(b, p0 - 1, p0 * p0), (b, p0)
            assert (out[2] << 64 | out[3]) == pow(b, p1 - 1, p1 * p1), (b, p1)


def test_radix28_any_base(emu):
    """Any base through the radix-2^28 ladder with small-constant multiplication, 2^50 <= p < 2^64."""
    ps = [(1 << 50) + 27, (1 << 63) - 25, (1 << 63) + 29, (1 << 64) - 59, (1 << 64) - 1, 10 ** 18 + 9] + \
        samples(50, 64, 800, 9)
    rng = random.Random(22)
    for b in BASES:
        for p in ps:
            if p < (1 << 50):
                continue
            eb = min(64, (p - 1).bit_length() + rng.randint(0, 2))
            assert call(emu.emu_m28c, b, p, eb) == pow(b, p - 1, p * p), (b, p)
