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
    L.emu_montb.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, P, P]
    L.emu_sprp2.argtypes = [ctypes.c_uint64]
    L.emu_psp_possible.argtypes = [ctypes.c_uint64]
    L.emu_m28.argtypes = [ctypes.c_uint64, ctypes.c_int, P, P]
    L.emu_fp64x2.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P, ctypes.c_int]
    L.emu_fp64c.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P, ctypes.c_int]
    L.emu_m28c.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, P, P]
    L.emu_psp_order.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8)]
    L.emu_fp64c2.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, P, ctypes.c_int]
    L.emu_neg_mod_fp.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    L.emu_neg_mod_fp.restype = ctypes.c_uint32
    L.emu_neg_mod_fp_random.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.emu_neg_mod_fp_random.restype = ctypes.c_uint64
    L.emu_recip_near.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    L.emu_recip_near.restype = ctypes.c_double
    L.emu_fp64x2_near.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, P]
    L.emu_m28_sqr.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    return L


def call(fn, *args):
    hi, lo = ctypes.c_uint64(), ctypes.c_uint64()
    fn(*args, ctypes.byref(hi), ctypes.byref(lo))
    return (hi.value << 64) | lo.value


F64_PMAX = 9 << 47   # the FP64 p-adic class bound of wf_arith.h


def top_of_class(n, seed):
    """Odd numbers from 2^50 to the FP64 class bound, with both ends."""
    rng = random.Random(seed)
    return [(1 << 50) + 1, (1 << 50) + 27, F64_PMAX - 1, F64_PMAX - 3, 12 * 10 ** 14 + 7] + \
        [rng.randrange(1 << 50, F64_PMAX) | 1 for _ in range(n)]


def samples(lo_bits, hi_bits, n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        bits = rng.randint(lo_bits, hi_bits)
        p = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        out.append(p)
    return out


def check_pair(out, b, p0, p1, eb):
    """Both ladders of a pair: the normalised residue equals b^(p-1) mod p^2, and the kernel's
    fast finish from the raw digits (f64_quotient) is taken exactly when r = 1 (mod p), with
    q = (r - 1) / p.  The samples mix primes and odd composites, so both branches run."""
    for k, p in ((0, p0), (1, p1)):
        r = out[4 * k] << 64 | out[4 * k + 1]
        assert r == pow(b, p - 1, p * p), (b, p, p0, p1, eb)
        assert out[4 * k + 2] == (r % p == 1), (b, p, eb)
        if out[4 * k + 2]:
            assert out[4 * k + 3] == (r - 1) // p, (b, p, eb)


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
    out = (ctypes.c_uint64 * 8)()
    for p0 in ps:
        if p0 < 3:
            continue
        p1 = rng.choice(ps)
        if p1 < 3:
            p1 = 3
        eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
        for e in (eb, min(50, eb + rng.randint(0, 7))):
            emu.emu_fp64x2(p0, p1, e, out, 0)
            check_pair(out, 2, p0, p1, e)


def test_fp64_two_ladders_big(emu):
    """The shorter window step for p >= 2^32 (segments at or above 2^32): pairs of any sizes
    from 2^32 + 15 to 2^50, the wavefront's exponent length or more."""
    ps = [4294967311, (1 << 32) + 61, (1 << 33) + 17, (1 << 50) - 27, (1 << 49) + 1, 10 ** 15 + 37] + \
        [p for p in samples(33, 50, 12000, 17)]
    rng = random.Random(13)
    out = (ctypes.c_uint64 * 8)()
    for p0 in ps:
        p1 = rng.choice(ps)
        eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
        for e in (eb, min(50, eb + rng.randint(0, 7))):
            emu.emu_fp64x2(p0, p1, e, out, 1)
            check_pair(out, 2, p0, p1, e)


def test_fp64_top_of_class(emu):
    """2^50 <= p < 9 * 2^47, where the quotient errors are larger than below 2^50 (the bounds
    hold over the five squarings between window steps): the single ladder, the two-ladder
    form with both window steps, paired with smaller big p."""
    ps = top_of_class(6000, 50)
    small = samples(33, 50, 200, 51)
    rng = random.Random(52)
    out = (ctypes.c_uint64 * 8)()
    for p0 in ps:
        assert call(emu.emu_fp64, p0) == pow(2, p0 - 1, p0 * p0), p0
        p1 = rng.choice(ps + small) | 1
        eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
        for big in (0, 1):
            emu.emu_fp64x2(p0, p1, eb, out, big)
            check_pair(out, 2, p0, p1, eb)


def test_fp64_bounds_to_the_class_top():
    """The error model of wf_arith.h's FP64 ladder, propagated in exact rationals: from the
    bounds a window step leaves (|a| <= p/2 (1 + 2^-20), |b| <= 1.1 p), five squarings at
    p = 9 * 2^47 keep every rounded quotient inside round_mul's range (|x y| < 2^51) and every
    digit below 2^53; at 1.25 * 2^50 the margin is smaller but still there, and the model
    fails before 2^51 (so it is not vacuous)."""
    from fractions import Fraction as Fr
    u = Fr(1, 1 << 53)
    rel = u + u * (1 + Fr(1, 128)) + u * u * (1 + Fr(1, 128))   # fl(x) * ip against x / p

    def worst(p):
        a, b, w = Fr(1, 2) * (1 + Fr(1, 1 << 20)), Fr(11, 10), Fr(0)
        for _ in range(5):
            # |h1 ip| <= a^2 p (1 + rel), |h2 ip| <= a b p (1 + rel), against 2^51
            w = max(w, a * a * p * (1 + rel) / (1 << 51), a * b * p * (1 + rel) / (1 << 51))
            k1err = Fr(1, 2) + a * a * p * rel          # |k1 - a^2/p|
            r = Fr(1, 2) + a * b * p * rel              # |r2| / p
            a, b = k1err, a * a + k1err / p + 2 * r
            w = max(w, b * p / (1 << 53))
        return w
    assert worst(Fr(1 << 50)) < Fr(56, 100)
    assert worst(Fr(F64_PMAX)) < Fr(69, 100)
    assert worst(Fr(5 << 48)) < 1
    assert worst(Fr(1 << 51)) > 1


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


@pytest.mark.parametrize("L,full,lo,hi", [(2, 0, 2, 31), (3, 0, 32, 47), (4, 0, 48, 63), (4, 1, 64, 64),
                                          (4, 1, 63, 64)])
def test_montgomery_any_base(emu, L, full, lo, hi):
    """The binary ladder of the 32-bit limb classes for bases other than 2, against Python's
    pow: bases up to 2^32 - 1, bases at or above p^2 (p = 3, 5, 7), and p dividing b."""
    ps = samples(lo, hi, 600, L * 10 + full + 100)
    if L == 4 and full:
        ps += [(1 << 64) - 1, (1 << 64) - 59, (1 << 63) + 1]
    if L == 2:
        ps += [3, 5, 7, 9, 11, 13, 65537, (1 << 31) - 1]
    rng = random.Random(L * 7 + full)
    bases = [3, 5, 6, 7, 10, 30, 97, 1625, 65537, 4294967291, 4294967295] + [rng.randrange(3, 1 << 32) for _ in range(4)]
    for p in ps:
        if p < 3:
            continue
        eb = (p - 1).bit_length()
        for b in bases:
            want = pow(b, p - 1, p * p)
            assert call(emu.emu_montb, L, full, b, p, eb) == want, (L, b, p)
            assert call(emu.emu_montb, L, full, b, p, min(64, eb + 2)) == want, (L, b, p)


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
    # every size, so that d's five leading bits (the FP64-estimated start) take every form
    ns += [rng.getrandbits(rng.randint(2, 64)) | 1 for _ in range(3000)]
    # long runs of squarings after the ladder (n - 1 = d 2^s, s up to 62), primes among them
    ns += [(rng.getrandbits(max(1, 63 - s)) << s) + 1 for s in range(1, 63) for _ in range(20)]
    ns += [3 * (1 << 30) + 1, 15 * (1 << 27) + 1, 18446744066047760377, 18446743208455367653]
    for n in ns:
        if n < 3 or n % 2 == 0 or n >= 1 << 64:
            continue
        assert bool(emu.emu_sprp2(n)) == strong2(n), n
    assert sum(strong2(n) for n in ns if 3 <= n < 1 << 64 and n % 2) > 2000   # many true cases


def test_radix28_montgomery(emu):
    ps = samples(47, 64, 8000, 99) + [(1 << 47) + 1, (1 << 50) - 27, (1 << 63) - 25, (1 << 63) + 29,
                                      (1 << 64) - 59, (1 << 64) - 1, (1 << 64) - 3, 10 ** 18 + 9]
    for p in ps:
        if p < (1 << 47):
            continue
        eb = (p - 1).bit_length()
        assert call(emu.emu_m28, p, eb) == pow(2, p - 1, p * p), p
        assert call(emu.emu_m28, p, min(64, eb + 2)) == pow(2, p - 1, p * p), p


def test_radix28_lazy_squaring_bounds(emu):
    """m28_sqr on the largest limbs the ladder can hand it: a value below 17.01 N with 28-bit
    limbs 0..3, doubled (limbs of 29 bits, below 34.02 N), and values up to 64 N.  The result
    must be x^2 / R mod N, below 17.01 N, with limbs 0..3 below 2^28: the bound that keeps the
    next squaring's column sums inside 64 bits and lets the final REDC reduce in one step."""
    R = 1 << 140
    M = (1 << 28) - 1
    rng = random.Random(28)
    ps = [(1 << 64) - 59, (1 << 64) - 1, (1 << 63) + 29, (1 << 50) + 27, (1 << 47) + 1] + samples(47, 64, 300, 5)
    for p in ps:
        N = p * p
        rinv = pow(R, -1, N)
        tops = [17 * N + N // 100, 34 * N + N // 50, 64 * N - 1]
        cases = []
        for top in tops:
            cases.append(top)
            cases += [rng.randrange(top) for _ in range(6)]
        for X in cases:
            if X >= 34 * N and X < 64 * N - 1 and rng.random() < 0.5:
                continue
            limbs = [(X >> (28 * i)) & M for i in range(4)] + [X >> 112]
            if X <= 34 * N + N // 50 and X % 2 == 0 and rng.random() < 0.5:
                # the doubled form the ladder really holds: each limb twice a 28-bit limb
                h = X // 2
                limbs = [2 * ((h >> (28 * i)) & M) for i in range(4)] + [2 * (h >> 112)]
            assert sum(v << (28 * i) for i, v in enumerate(limbs)) == X
            xin = (ctypes.c_uint32 * 5)(*limbs)
            xout = (ctypes.c_uint32 * 5)()
            emu.emu_m28_sqr(p, xin, xout)
            Y = sum(xout[i] << (28 * i) for i in range(5))
            assert all(xout[i] <= M for i in range(4)), (p, X)
            assert Y % N == X * X * rinv % N, (p, X)
            assert Y < 17.01 * N, (p, X, Y / N)


def test_psp_filter_from_tile_residues(emu):
    """k_psp forms psp_possible's residues from the tile's first number plus the offset; the
    verdicts equal psp_possible(n) on whole tiles, including tiles whose residues wrap."""
    emu.emu_psp_tile_mismatch.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    emu.emu_psp_tile_mismatch.restype = ctypes.c_uint32
    rng = random.Random(97)
    starts = [3, (1 << 64) - 4097, 18446744066047760377 - 2 * 100, 2047 - 20] + \
        [rng.getrandbits(64) | 1 for _ in range(40)] + [rng.getrandbits(40) | 1 for _ in range(20)]
    # tiles whose residues wrap modulo each product inside the tile
    for M in (334639305, 58642669, 600662303, 3024658859, 97):
        base = rng.getrandbits(60) * M
        starts.append((base - 2 * 777) | 1)
    for A in starts:
        if A + 2 * 2048 >= 1 << 64:
            A = (1 << 64) - 2 * 2048 - 1
        assert emu.emu_psp_tile_mismatch(A, 2048) == 0, A


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


def _order2(r):
    k, x = 1, 2 % r
    while x != 1:
        x = x * 2 % r
        k += 1
    return k


def _v2(x):
    return (x & -x).bit_length() - 1


def psp_possible_definition(n):
    """For every prime r < 100 dividing n: ord_r(2) divides n - 1, and the 2-adic valuations
    of these orders are all equal."""
    vs = set()
    for r in (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97):
        if n % r == 0:
            o = _order2(r)
            if (n - 1) % o:
                return False
            vs.add(_v2(o))
    return len(vs) <= 1


def test_psp_filter_equals_its_definition(emu):
    """The branch-free filter against the condition written out with Python integers, on
    random odd numbers up to 2^64 and on products of small primes (which exercise many
    factors and every valuation at once)."""
    rng = random.Random(11)
    small = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
    ns = [rng.randrange(3, 1 << 64) | 1 for _ in range(40000)] + list(range(3, 60000, 2))
    for _ in range(40000):
        m = 1
        for r in rng.sample(small, rng.randint(1, 4)):
            m *= r ** rng.randint(1, 2)
        k = rng.randrange(1, max(2, ((1 << 64) - 1) // m))
        n = m * k
        if n % 2 == 0:
            n += m if m % 2 else 0
        if n % 2 and 3 <= n < 1 << 64:
            ns.append(n)
    bad = [n for n in ns if bool(emu.emu_psp_possible(n)) != psp_possible_definition(n)]
    assert bad == []


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
    out = (ctypes.c_uint64 * 8)()
    for b in BASES:
        for p0 in ps:
            if p0 < 3:
                continue
            p1 = rng.choice(ps) | 1
            if p1 < 3:
                p1 = 3
            eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
            eb = min(50, eb + rng.randint(0, 3))
            emu.emu_fp64c(b, p0, p1, eb, out, 0)
            check_pair(out, b, p0, p1, eb)
            if p0 >= 1 << 32 and p1 >= 1 << 32:           # the shorter window step
                emu.emu_fp64c(b, p0, p1, eb, out, 1)
                check_pair(out, b, p0, p1, eb)


def test_fp64_any_base_big(emu):
    """The shorter window step of any base for p >= 2^32, from 2^32 + 15 to 2^50."""
    rng = random.Random(29)
    ps = [4294967311, (1 << 32) + 61, (1 << 50) - 27] + samples(33, 50, 600, 31)
    out = (ctypes.c_uint64 * 8)()
    for b in BASES:
        for p0 in ps:
            p1 = rng.choice(ps)
            eb = min(50, max((p0 - 1).bit_length(), (p1 - 1).bit_length()) + rng.randint(0, 3))
            emu.emu_fp64c(b, p0, p1, eb, out, 1)
            check_pair(out, b, p0, p1, eb)


def test_fp64_any_base_top_of_class(emu):
    """Every base through the FP64 two-ladder form with both window steps, 2^50 <= p < 9 * 2^47."""
    rng = random.Random(53)
    ps = top_of_class(150, 54)
    out = (ctypes.c_uint64 * 8)()
    for b in BASES:
        for p0 in ps:
            p1 = rng.choice(ps)
            eb = max((p0 - 1).bit_length(), (p1 - 1).bit_length())
            for big in (0, 1):
                emu.emu_fp64c(b, p0, p1, eb, out, big)
                check_pair(out, b, p0, p1, eb)


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


# ---- strong Lucas (Selfridge) and BPSW --------------------------------------------------

def _lucas_uv_matrix(P, Q, k, n):
    """U_k, V_k mod n from the 2x2 matrix [[P, -Q], [1, 0]]^k = [[U_(k+1), -Q U_k], [U_k, -Q U_(k-1)]],
    V_k = 2 U_(k+1) - P U_k: the definition, by a route unlike the kernel's V ladder."""
    def mul(A, B):
        return [[(A[0][0] * B[0][0] + A[0][1] * B[1][0]) % n, (A[0][0] * B[0][1] + A[0][1] * B[1][1]) % n],
                [(A[1][0] * B[0][0] + A[1][1] * B[1][0]) % n, (A[1][0] * B[0][1] + A[1][1] * B[1][1]) % n]]
    R = [[1, 0], [0, 1]]
    M = [[P % n, (-Q) % n], [1, 0]]
    while k:
        if k & 1:
            R = mul(R, M)
        M = mul(M, M)
        k >>= 1
    u1, u = R[0][0], R[1][0]
    return u, (2 * u1 - P * u) % n


def _jacobi(a, n):
    a %= n
    t = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                t = -t
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            t = -t
        a %= n
    return t if n == 1 else 0


def _slprp_def(n):
    r = int(n ** 0.5)
    while r * r > n:
        r -= 1
    while (r + 1) ** 2 <= n:
        r += 1
    if r * r == n:
        return False
    D = 5
    while True:
        j = _jacobi(D, n)
        if j == -1:
            break
        if j == 0 and abs(D) != n:
            return False
        D = -(D + 2) if D > 0 else 2 - D
    Q = (1 - D) // 4
    d, s = n + 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    u, v = _lucas_uv_matrix(1, Q, d, n)
    if u == 0 or v == 0:
        return True
    return any(_lucas_uv_matrix(1, Q, d << r, n)[1] == 0 for r in range(1, s))


@pytest.fixture(scope="module")
def emu_lucas(emu):
    emu.emu_slprp.argtypes = [ctypes.c_uint64]
    emu.emu_bpsw.argtypes = [ctypes.c_uint64]
    emu.emu_bpsw.restype = ctypes.c_uint32
    emu.emu_jacobi.argtypes = [ctypes.c_int, ctypes.c_uint64]
    return emu


def test_jacobi_small(emu_lucas):
    from sympy import jacobi_symbol
    rng = random.Random(31)
    ns = [rng.getrandbits(64) | 1 for _ in range(400)] + [3, 5, 7, 9, 15, 21, 25, 5459, (1 << 64) - 1]
    for n in ns:
        for D in (5, -7, 9, -11, 13, -15, 17, -19, 21, -23, 25, -101, 1001):
            assert emu_lucas.emu_jacobi(D, n) == jacobi_symbol(D, n), (D, n)


def test_strong_lucas_small_exhaustive(emu_lucas):
    """Every odd n < 30000 against the matrix-power definition; the composites that pass are
    the first strong Lucas pseudoprimes (OEIS A217255)."""
    passing_composites = []
    for n in range(3, 30000, 2):
        got = bool(emu_lucas.emu_slprp(n))
        assert got == _slprp_def(n), n
        if got and not _isprime(n):
            passing_composites.append(n)
    assert passing_composites == [5459, 5777, 10877, 16109, 18971, 22499, 24569, 25199]


def test_strong_lucas_64bit(emu_lucas):
    """Large n against sympy's is_strong_lucas_prp (a library routine) and the matrix
    definition: random odd n, primes near 2^64, squares, semiprimes with large factors."""
    from sympy.ntheory.primetest import is_strong_lucas_prp
    rng = random.Random(37)
    ns = [rng.getrandbits(64) | 1 for _ in range(1500)] + [rng.getrandbits(rng.randint(20, 63)) | 1 for _ in range(1500)]
    ns += [(1 << 64) - 59, (1 << 64) - 83, (1 << 64) - 1, (1 << 63) + 29, 4294967291 ** 2, 4294967279 * 4294967291,
           3825123056546413051, 1000000007 * 998244353]
    ns += [q for q in (rng.randrange(1 << 62, 1 << 64) | 1 for _ in range(4000)) if _isprime(q)][:300]
    for n in ns:
        want = is_strong_lucas_prp(n)
        assert bool(emu_lucas.emu_slprp(n)) == want, n
        if n < (1 << 40):
            assert _slprp_def(n) == want, n


def test_bpsw_flags(emu_lucas):
    """bit 1: base-2 strong probable prime, bit 2: strong Lucas probable prime, bit 4: prime.
    Base-2 strong pseudoprimes set only bit 1, strong Lucas pseudoprimes only bit 2, and the
    verdict agrees with sympy.isprime on random 64-bit numbers and on every n < 20000."""
    import sympy
    for n in (2047, 3277, 4033, 4681, 8321, 3825123056546413051, 18446744066047760377):
        assert emu_lucas.emu_bpsw(n) == 1, n
    for n in (5459, 5777, 10877, 16109, 18971):
        assert emu_lucas.emu_bpsw(n) == 2, n
    assert [emu_lucas.emu_bpsw(n) for n in (0, 1, 2, 3, 4, 5, 9)] == [0, 0, 4, 7, 0, 7, 0]
    for n in range(0, 20000):
        assert bool(emu_lucas.emu_bpsw(n) & 4) == sympy.isprime(n), n
    rng = random.Random(41)
    for _ in range(3000):
        n = rng.getrandbits(64) | 1
        assert bool(emu_lucas.emu_bpsw(n) & 4) == sympy.isprime(n), n


def test_psp_order_sieve(emu):
    """k_psp's order sieve against its definition: an odd n is ruled out exactly when some
    prime 101 <= p < 16384 divides it with ord_p(2) not dividing n - 1 (orders from sympy's
    n_order, multiples found by plain division).  Tiles near 2^64, 2^50 and small ones where
    n = p itself appears (never ruled out); every base-2 Fermat pseudoprime the tiles hold,
    and the listed strong pseudoprimes, are kept."""
    from sympy import n_order, primerange
    PMAX, NODD = 16384, 2048
    orders = [(p, n_order(2, p)) for p in primerange(101, PMAX)]
    rng = random.Random(41)
    tiles = [1, 101, 16383, 3 * 10 ** 9 + 1] + [(1 << 64) - 2 * NODD - 1, (1 << 64) - 4095] + \
        [rng.randrange(1, 1 << 64) | 1 for _ in range(6)] + [rng.randrange(1 << 49, 1 << 50) | 1 for _ in range(4)]
    tiles += [n - 2 * rng.randrange(NODD) for n in (2047, 3277, 4681, 15841, 29341, 52633, 3825123056546413051,
                                                     18446744066047760377)]
    out = (ctypes.c_uint8 * NODD)()
    for A in tiles:
        A = max(A, 1) | 1
        A = min(A, (1 << 64) - 2 * NODD + 1)
        emu.emu_psp_order(A, PMAX, NODD, out)
        want = [0] * NODD
        for p, L in orders:
            n = -(-A // p) * p
            while n < A + 2 * NODD:
                if n % 2 and (n - 1) % L:
                    want[(n - A) // 2] = 1
                n += p
        assert list(out) == want, A
        for i in range(NODD):
            n = A + 2 * i
            if out[i] and n > 1:
                assert pow(2, n - 1, n) != 1, n


def test_fp64_two_bases_one_prime(emu):
    """The composite-base mode's ladder pairs: two bases on one prime with the smaller window
    of the two (so 2 and 3 run at 4 bits, 29 at 2), p from 3 to 2^50, and the shorter window
    step above 2^32."""
    rng = random.Random(43)
    pairs = [(2, 3), (5, 7), (11, 13), (17, 19), (23, 23), (2, 29), (3, 65521), ((1 << 31) - 1, 2), (6, 10)]
    ps = [3, 5, 7, 11, 13, 29, 31, 1093, 3511, (1 << 50) - 27] + samples(2, 50, 800, 44)
    out = (ctypes.c_uint64 * 8)()
    for A, B in pairs:
        for p in ps:
            if p < 3:
                continue
            eb = min(50, (p - 1).bit_length() + rng.randint(0, 3))
            for big in ([0, 1] if p >= 1 << 32 else [0]):
                emu.emu_fp64c2(A, B, p, eb, out, big)
                for k, b in ((0, A), (1, B)):
                    r = out[4 * k] << 64 | out[4 * k + 1]
                    assert r == pow(b, p - 1, p * p), (A, B, b, p, eb, big)
                    assert out[4 * k + 2] == (r % p == 1), (b, p)
                    if out[4 * k + 2]:
                        assert out[4 * k + 3] == (r - 1) // p, (b, p)



def test_neg_mod_fp_edges(emu):
    """The sieve's and fill's one-sided FP64 placement: (-base) mod q for 512 <= q < 2^31 and
    base < 2^53.  The quotient estimate is biased low, so the cases that would break a wrong
    bias are base / q at or just around an integer, at both ends of the base range and for the
    smallest and largest q."""
    top = (1 << 53) - 1
    qs = [512, 521, 1021, 65521, 1048573, (1 << 24) - 3, 94906249, (1 << 31) - 1]
    for q in qs:
        for m in (0, 1, 2, 12345, top // q - 1, top // q):
            for d in (0, 1, 2, q // 256, q // 128, q // 128 + 1, q // 64, q // 2, q - 2, q - 1):
                base = m * q + d
                if base > top:
                    continue
                assert emu.emu_neg_mod_fp(base, q) == (-base) % q, (base, q)
        for base in (0, 1, q - 1, q, q + 1, top, top - 1, top - q, 10 ** 15, 10 ** 15 + 10 ** 11):
            assert emu.emu_neg_mod_fp(base, q) == (-base) % q, (base, q)


def test_neg_mod_fp_random(emu):
    """Four million random (base, q) pairs, half of them with base / q within 2^-6 of an
    integer, against exact integer arithmetic (inside the harness)."""
    assert emu.emu_neg_mod_fp_random(4_000_000, 0x243F6A8885A308D3) == 0


def test_recip_near_bound(emu):
    """f64_recip_near: 1/p from the segment's 1/base (base >= 2^40, offset f < 2^22) within
    2^-53 (1 + 2^-7) of 1/p, relative, checked exactly with rationals; the edges are the
    smallest base, the largest offset and p just below the FP64 class bound 9 * 2^47."""
    from fractions import Fraction
    rng = random.Random(7)
    cases = [((1 << 40), (1 << 22) - 1), ((1 << 40), 1), ((1 << 50) - (1 << 22), (1 << 22) - 1),
             (10 ** 15, 0), (10 ** 15, 1 << 21), (30 * ((1 << 45) // 30), 12345),
             (30 * ((F64_PMAX - (1 << 22)) // 30), (1 << 22) - 1)]
    for _ in range(20000):
        base = 30 * (rng.randrange(1 << 40, F64_PMAX - (1 << 22)) // 30)
        if base < 1 << 40:
            base += 30
        cases.append((base, rng.randrange(0, 1 << 22)))
    bound = Fraction(1, 1 << 53) * (1 + Fraction(1, 128))
    exact = 0
    for base, f in cases:
        p = base + f
        ip = emu.emu_recip_near(base, f)
        err = abs(Fraction(ip) * p - 1)
        assert err <= bound, (base, f, float(err))
        exact += ip == 1.0 / p
    assert exact > 0.9 * len(cases)      # informational: mostly the correctly rounded value


def test_fp64_two_ladders_near(emu):
    """The kernel's base-2 ladder pair above 2^40 with 1/p from f64_recip_near, on primes and
    odd composites of segments at 2^40, 1e15, just below 2^50, at 1.2e15 and just below the
    FP64 class bound, against Python's pow."""
    rng = random.Random(11)
    out = (ctypes.c_uint64 * 8)()
    for base in (30 * ((1 << 40) // 30 + 1), 10 ** 15, 30 * (((1 << 50) - (1 << 22)) // 30), 12 * 10 ** 14,
                 30 * ((F64_PMAX - (1 << 22)) // 30)):
        eb = (base + (1 << 22)).bit_length()
        for _ in range(300):
            f0, f1 = rng.randrange(1, 1 << 22, 2), rng.randrange(1, 1 << 22, 2)
            emu.emu_fp64x2_near(base, f0, f1, eb, out)
            check_pair(out, 2, base + f0, base + f1, eb)
