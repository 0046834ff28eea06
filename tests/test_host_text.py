"""The host program's text layer (mathpub_amd/host/wf_text.c) on the CPU: the number and
base-list grammar, the checkpoint file, the chunk-log line and SUMMARY, byte for byte in the
formats of A001220/wieferich.c (grammar :329-416 and :816-849, checkpoint :1211-1316,
chunk log :1166-1172, SUMMARY :1505-1523); then the same code, the whole wf command's
CPU-only paths and the oracle built with -fsanitize=address,undefined and fed malformed
input."""
import os
import random
import re
import subprocess
from fractions import Fraction

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mathpub_amd", "host")
DRIVER_SRC = os.path.join(ROOT, "tests", "host", "text_driver.c")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
SAN_ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:exitcode=99:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
TWO64 = 1 << 64


def _gcc(out, srcs, extra=()):
    subprocess.check_call(["gcc", "-O1", "-g", "-std=gnu11", "-Wall", "-Wextra", "-Werror", "-I" + HOST,
                           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle")] + list(extra) +
                          ["-o", str(out)] + srcs + ["-lm", "-pthread"])
    return str(out)


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    d = tmp_path_factory.mktemp("text")
    srcs = [DRIVER_SRC, os.path.join(HOST, "wf_text.c")]
    return _gcc(d / "plain", srcs), _gcc(d / "san", srcs, SAN)


def call(exe, *args, stdin=None, env=None):
    p = subprocess.run([exe] + [str(a) for a in args], input=stdin, capture_output=True, encoding="utf-8",
                       errors="replace", env=env, timeout=120)
    return p.returncode, p.stdout, p.stderr


def clean(rc, err):
    """No sanitizer report and a normal exit."""
    assert rc == 0 and "runtime error:" not in err and "AddressSanitizer" not in err and "LeakSanitizer" not in err, err


# ---- numbers ----

NUMBERS = [  # (text, result) from the grammar's definition and the reference's own examples
    ("1000", "ok 1000"), ("1e15", "ok 1000000000000000"), ("1.2e15", "ok 1200000000000000"), ("5.e3", "ok 5000"),
    (".5e1", "ok 5"), ("1.50e3", "ok 1500"), ("2^63", "ok 9223372036854775808"), ("2^64", "ok 18446744073709551616"),
    ("2^63-25", "ok 9223372036854775783"), ("1e15+4e6", "ok 1000000004000000"), ("0^0", "ok 1"), ("0^7", "ok 0"),
    ("1^99999999999", "ok 1"), ("0e5", "ok 0"), ("18446744073709551616", "ok 18446744073709551616"),
    ("2^64-1", "ok 18446744073709551615"), ("1E3", "ok 1000"),
    ("1e-5", "bad"), ("1e+5", "bad"), ("e5", "bad"), (".e5", "bad"), ("1.5e0", "bad"), ("2^65", "bad"),
    ("18446744073709551617", "bad"), ("1-2-3", "bad"), ("-5", "bad"), ("5-", "bad"), ("+5", "bad"), ("", "bad"),
    ("1.2.3e4", "bad"), ("1e5e3", "bad"), ("2^3^2", "bad"), ("1e3^2", "bad"), ("12a", "bad"), (" 5", "bad"),
    ("1" * 64, "bad"), ("0" * 62 + "7", "ok 7"),
    ("18446744073709551616+1", "overflow"), ("2^64+2^64", "overflow"), ("5-7", "negative"), ("0-1", "negative"),
]


@pytest.mark.parametrize("text,want", NUMBERS)
def test_number_examples(drivers, text, want):
    assert call(drivers[0], "number", text)[1].strip() == want


def _digits(t):
    return re.fullmatch(r"[0-9]+", t) is not None


def model_term(t):
    """A term's value from the grammar's definition, or None."""
    if not t or len(t) > 63:
        return None
    if "^" in t:
        b, _, e = t.partition("^")
        if not (_digits(b) and _digits(e)):
            return None
        b, e = int(b), int(e)
        if b > TWO64 or e > TWO64:
            return None
        if b <= 1:
            return 1 if (e == 0 or b == 1) else 0
        if e > 64:
            return None
        v = b ** e
        return v if v <= TWO64 else None
    m = re.search("[eE]", t)
    if m:
        mant, ex = t[:m.start()], t[m.start() + 1:]
        if not _digits(ex) or int(ex) > TWO64:
            return None
        if not re.fullmatch(r"[0-9]*\.?[0-9]*", mant) or not re.search("[0-9]", mant):
            return None
        ip, _, fp = mant.partition(".")
        v = Fraction(int(ip or "0")) + (Fraction(int(fp), 10 ** len(fp)) if fp else 0)
        if v == 0:
            return 0
        if int(ex) > 40:
            return None
        v *= 10 ** int(ex)
        return int(v) if v.denominator == 1 and v <= TWO64 else None
    return int(t) if _digits(t) and int(t) <= TWO64 else None


def model_number(s):
    op = 0
    for i in range(1, len(s)):
        if s[i] in "+-" and s[i - 1] not in "eE":
            op = i
    if not op:
        v = model_term(s)
        return "bad" if v is None else "ok %d" % v
    a, b = model_term(s[:op]), model_term(s[op + 1:])
    if a is None or b is None:
        return "bad"
    if s[op] == "+":
        return "overflow" if a + b > TWO64 else "ok %d" % (a + b)
    return "negative" if b > a else "ok %d" % (a - b)


def test_number_model_and_fuzz(drivers):
    """Random strings over the grammar's alphabet: the C parser agrees with the definition,
    and the sanitized build reports nothing."""
    rng = random.Random(5)
    alpha = "0123456789^eE.+-" * 3 + "x "
    corpus = [t for t, _ in NUMBERS]
    for _ in range(6000):
        corpus.append("".join(rng.choice(alpha) for _ in range(rng.randint(1, 24))))
    for _ in range(2000):   # well-formed shapes with random digits
        d = lambda: str(rng.randrange(10 ** rng.randint(1, 21)))
        corpus.append(rng.choice([d() + "e" + str(rng.randint(0, 25)), d() + "." + d() + "e" + str(rng.randint(0, 30)),
                                  str(rng.randint(0, 20)) + "^" + str(rng.randint(0, 70)),
                                  d() + rng.choice("+-") + d()]))
    text = "\n".join(corpus) + "\n"
    rc, out, err = call(drivers[0], "numbers", stdin=text)
    assert rc == 0
    got = out.splitlines()
    assert len(got) == len(corpus)
    for s, g in zip(corpus, got):
        assert g == model_number(s), s
    rc, out2, err = call(drivers[1], "numbers", stdin=text, env=SAN_ENV)
    clean(rc, err)
    assert out2 == out


# ---- base lists ----

BASES = [
    ("oeis", "ok 3,5,6,7,10,12,13,14,15,17,18,19,20,22,23,26,30"),
    ("known", "ok 2,3,5,6,7,10,11,12,13,14,15,17,18,19,20,21,22,23,24,26,28,29,30"),
    ("2", "ok 2"), ("30,2,3,2", "ok 2,3,30"), ("2-6", "ok 2,3,4,5,6"), ("2^3-10", "ok 8,9,10"),
    (",,3,,2,", "ok 2,3"), ("4294967295", "ok 4294967295"),
    ("2^32-1", "err bad base or range '2^32-1'"),   # a range splits at its first '-': 2^32 .. 1
    ("2-5,oeis", "ok 2,3,4,5,6,7,10,12,13,14,15,17,18,19,20,22,23,26,30"),
    ("1", "err bad base or range '1'"), ("5-3", "err bad base or range '5-3'"), ("2^32", "err bad base or range '2^32'"),
    ("", "err between 1 and 64 bases are supported"), (",", "err between 1 and 64 bases are supported"),
    ("2-66", "err too many bases (max 64)"), ("x", "err bad number 'x'"), ("2-x", "err bad number 'x'"),
    ("3," * 600, "err base list too long"),
]


@pytest.mark.parametrize("text,want", BASES)
def test_base_lists(drivers, text, want):
    assert call(drivers[0], "bases", text)[1].strip() == want


def test_base_list_fuzz_sanitized(drivers):
    rng = random.Random(9)
    alpha = "0123456789,-^e" * 4 + "oeisknw"
    corpus = ["".join(rng.choice(alpha) for _ in range(rng.randint(0, 40))) for _ in range(3000)]
    corpus += ["2-4294967295", "4294967290-4294967295", "0-0", "2-" + "9" * 30, "oeis,known," * 50]
    rc, out, err = call(drivers[1], "bases-list", stdin="\n".join(corpus) + "\n", env=SAN_ENV)
    clean(rc, err)
    for line in out.splitlines():
        if line.startswith("ok "):
            b = [int(x) for x in line[3:].split(",")]
            assert b == sorted(set(b)) and 1 <= len(b) <= 64 and b[0] >= 2 and b[-1] < 2 ** 32


# ---- checkpoint ----

CKPT = """wieferich checkpoint 1
bases 2,3,5
start 1000000000000000
chunk 1000000000
end 1000000010000000
frontier 4
primes 123456789
tests 370370367
near 7
flt 0
cks 2 0123456789abcdef
cks 3 fedcba9876543210
cks 5 0000000000000001
hit 3 1000000000123 2
hit 5 1000000002001 3
end
"""


def test_checkpoint_round_trip(drivers, tmp_path):
    """A checkpoint in the reference's layout is read field by field and written back
    identically; a solution at or past the frontier is not written; END may be 2^64."""
    f = tmp_path / "st"
    f.write_text(CKPT)
    rc, out, err = call(drivers[0], "ckpt-rewrite", f)
    assert rc == 0 and out == CKPT
    rc, out, err = call(drivers[0], "ckpt", f)
    assert out.splitlines()[:4] == ["bases 2,3,5", "start 1000000000000000", "chunk 1000000000", "end 1000000010000000"]
    f.write_text(CKPT.replace("hit 5 1000000002001 3\n", "hit 5 1000000002001 3\nhit 7 99 4\n")
                 .replace("end 1000000010000000", "end 18446744073709551616"))
    rc, out, err = call(drivers[0], "ckpt-rewrite", f)
    assert out == CKPT.replace("end 1000000010000000", "end 18446744073709551616")
    # unknown lines are ignored, CRLF line ends are accepted
    f.write_text(CKPT.replace("near 7\n", "near 7\nfuture_field 12\n").replace("\n", "\r\n"))
    assert call(drivers[0], "ckpt-rewrite", f)[1] == CKPT


@pytest.mark.parametrize("text,msg", [
    ("", "is not a wieferich checkpoint"), ("hello\n", "is not a wieferich checkpoint"),
    (CKPT[:-4], "is truncated"), (CKPT.split("cks 2")[0], "is truncated"),
    (CKPT.replace("cks 2 0123456789abcdef", "cks 2 0123456789abcdeg"), ": bad cks line"),
    (CKPT.replace("cks 2 0123456789abcdef", "cks 2"), ": bad cks line"),
    (CKPT.replace("cks 2 0123456789abcdef", "cks 2 10123456789abcdef"), ": bad cks line"),
    (CKPT.replace("cks 2 0123456789abcdef", "cks 4294967296 0123456789abcdef"), ": bad cks line"),
    (CKPT.replace("hit 3 1000000000123 2", "hit 3 1000000000123"), ": bad hit line"),
    (CKPT.replace("hit 3 1000000000123 2", "hit 3 -1 2"), ": bad hit line"),
    (CKPT.replace("start 1000000000000000", "start 1e15"), ": bad start line"),
    (CKPT.replace("primes 123456789", "primes 18446744073709551616"), ": bad primes line"),
    (CKPT.replace("frontier 4", "frontier"), ": bad frontier line"),
    (CKPT.replace("end 1000000010000000", "end 18446744073709551617"), ": bad end line"),
])
def test_checkpoint_rejections(drivers, tmp_path, text, msg):
    f = tmp_path / "st"
    f.write_text(text)
    rc, out, err = call(drivers[0], "ckpt", f)
    assert rc == 0 and out.startswith("err %s" % f) and msg in out, out


def _mutants(rng):
    base = CKPT.encode()
    out = [base[:k] for k in range(len(base))]
    for _ in range(300):
        b = bytearray(base)
        for _ in range(rng.randint(1, 6)):
            b[rng.randrange(len(b))] = rng.randrange(256)
        out.append(bytes(b))
    lines = CKPT.splitlines(True)
    for _ in range(100):
        ls = list(lines)
        for _ in range(rng.randint(1, 4)):
            i = rng.randrange(len(ls))
            ls.insert(rng.randrange(len(ls)), ls[i]) if rng.random() < 0.5 else ls.pop(i)
        out.append("".join(ls).encode())
    out.append(base.replace(b"bases 2,3,5", b"bases " + b"7," * 60000))
    out.append(base.replace(b"near 7", b"near 7 " + b"x" * 300000))
    out.append(base.replace(b"flt 0\n", b"flt 0\n" + b"".join(b"cks %d 00\n" % k for k in range(70))))
    out.append(base.replace(b"hit 3", b"".join(b"hit 3 %d 1\n" % k for k in range(5000)) + b"hit 3"))
    out.append(base.replace(b"\n", b"\x00\n"))
    return out


def test_checkpoint_fuzz_sanitized(drivers, tmp_path):
    """Truncated, corrupted, reordered and oversized checkpoints: the sanitized reader either
    accepts the file or rejects it with a message, and never misbehaves."""
    rng = random.Random(3)
    f = tmp_path / "st"
    accepted = 0
    for m in _mutants(rng):
        f.write_bytes(m)
        rc, out, err = call(drivers[1], "ckpt", f, env=SAN_ENV)
        clean(rc, err)
        assert out.startswith("err ") or out.startswith("bases "), out[:200]
        accepted += out.startswith("bases ")
    assert accepted > 10


# ---- chunk log, SUMMARY, durations ----

def test_chunk_line(drivers):
    assert call(drivers[0], "chunk", "1000000000000000", "1000000100000000", "2895", "2:0123456789abcdef",
                "30:00000000000000ff")[1] == "C 1000000000000000 1000000100000000 2895 2:0123456789abcdef 30:00000000000000ff\n"
    assert call(drivers[0], "chunk", "18446744073709551516", "2^64", "1", "2:8000000000000000")[1] == \
        "C 18446744073709551516 18446744073709551616 1 2:8000000000000000\n"


def test_summary_block(drivers):
    out = call(drivers[0], "summary", "0", "100000", "9592", "9591", "0", "100", "5", "5", "0",
               "2:0123456789abcdef:11", "17:fedcba9876543210:22", "31:0000000000000001:33",
               "hit:2:1093:0", "hit:2:3511:1", "hit:17:46021:5")[1]
    assert out == ("SUMMARY [0, 100000) primes 9592 tests 9591 flt_errors 0 near(|A|<=100) 5\n"
                   "base   2 A001220  cks 0123456789abcdef  solutions: 1093 3511\n"
                   "base  17 A128668  cks fedcba9876543210  solutions: none\n"
                   "base  31          cks 0000000000000001  solutions: none\n")
    out = call(drivers[0], "summary", "7", "2^64", "1", "1", "1", "0", "0", "1", "1", "2:00000000000000aa:00000000000000bb")[1]
    assert out == ("SUMMARY [7, 18446744073709551616) primes 1 tests 1 flt_errors 1\n"
                   "base   2 A001220  cks 00000000000000aa  solutions: none\n"
                   "sumq base 2 00000000000000bb\n")


@pytest.mark.parametrize("s,want", [(0, "0.0s"), (12.34, "12.3s"), (245, "4m05s"), (7384, "2h03m04s"),
                                    (273900, "3d04h05m"), (-1, "?"), (2e12, "?"), (float("nan"), "?")])
def test_durations(drivers, s, want):
    assert call(drivers[0], "duration", repr(s))[1].strip() == want


# ---- the whole host and the oracle under the sanitizers ----

@pytest.fixture(scope="module")
def wf_san(tmp_path_factory):
    lib = os.path.join(ROOT, "mathpub_amd", "libwf.so")
    if not os.path.exists(lib):
        pytest.skip("libwf.so not built")
    from mathpub_amd import build
    out = str(tmp_path_factory.mktemp("wfsan") / "wf")
    build.build_host(out=out, extra=SAN + ["-g", "-O1"], rpath=os.path.dirname(lib))
    return out


@pytest.mark.parametrize("args,rc_want,msg", [
    ((), 2, "usage"), (("scan",), 2, "usage"), (("scan", "10", "-z"), 2, "usage"),
    (("scan", "1e-5", "2"), 1, "bad number '1e-5'"), (("scan", "0", "2^64+1"), 1, "overflow in"),
    (("scan", "5-7"), 1, "negative result in '5-7'"), (("scan", "10", "5"), 1, "END must be larger than START"),
    (("scan", "2^64-10", "2^64", "-c", "7"), 1, "END must be <= 2^64"), (("scan", "10", "-b", "1"), 1, "bad base"),
    (("scan", "10", "-b", "3," * 600), 1, "base list too long"), (("scan", "10", "-i", "0"), 1, "interval must be"),
    (("scan", "10", "-c", "0"), 1, "CHUNK must be positive"), (("check", "4"), 1, "P must be odd"),
    (("stats", "2^64"), 1, "END must be <= 2^64 - 2^20"), (("dcheck", "x", "2", "0"), 1, "STEP must be positive"),
    (("scan", "10", "--prove", "maybe"), 1, "--prove takes auto, off or on"),
])
def test_host_cli_errors_sanitized(wf_san, args, rc_want, msg):
    rc, out, err = call(wf_san, *args, env=SAN_ENV)
    assert rc == rc_want and msg in err, (rc, err)
    assert "runtime error:" not in err and "Sanitizer" not in err


def test_host_checkpoint_errors_sanitized(wf_san, tmp_path):
    """The scan command's checkpoint checks run before the GPU is opened: every malformed or
    mismatched checkpoint is rejected with its message, under the sanitizers."""
    f = tmp_path / "st"
    good = CKPT.replace("bases 2,3,5", "bases 2").replace("cks 3 fedcba9876543210\ncks 5 0000000000000001\n", "") \
        .replace("hit 3 1000000000123 2\nhit 5 1000000002001 3\n", "")
    cases = [(good.replace("start 1000000000000000", "start 999"), "was written for start=999 chunk=1000000000"),
             (good.replace("chunk 1000000000\n", "chunk 50\n"), "chunk=50, not start="),
             (good.replace("bases 2", "bases 3"), "was written for bases 3, not 2"),
             (good.replace("cks 2 0123456789abcdef\n", ""), "has no checksum for base 2"),
             (good.replace("cks 2 0123456789abcdef\n", "cks 2 0123456789abcdef\ncks 7 00\n"),
              "base 7 is not in the current base list"),
             (good.replace("frontier 4", "frontier 11"), "is past the new END (frontier 11 > 10 chunks)"),
             (good[:-4], "is truncated"), ("hello\n", "is not a wieferich checkpoint")]
    rng = random.Random(1)
    cases += [(m.decode("latin-1"), None) for m in _mutants(rng)[::7]]
    for text, msg in cases:
        f.write_bytes(text.encode("latin-1"))
        rc, out, err = call(wf_san, "scan", "1e15", "1e15+1e10", "-c", "1e9", "-b", "2", "-S", f, "-q", env=SAN_ENV)
        assert "runtime error:" not in err and "Sanitizer" not in err, err
        if msg is not None:
            assert rc == 1 and msg in err, (msg, err)


@pytest.fixture(scope="module")
def oracle_san(tmp_path_factory):
    d = tmp_path_factory.mktemp("osan")
    return _gcc(d / "oracle_san", [os.path.join(ROOT, "tests", "host", "oracle_driver.c"),
                                   os.path.join(ROOT, "oracle", "wf_oracle.c")], SAN)


@pytest.mark.parametrize("lo,hi,nearT,psp,bases", [
    (0, 10, 0, 0, [2]), (2, 3, 100, 0, [2, 5, 13]), (3, 2 * 10 ** 6, 100, 1, [2]),
    (10 ** 15, 10 ** 15 + 3 * 10 ** 5, 1000, 1, [2, 3, 30, 4294967291]), (1 << 40, (1 << 40) + 10 ** 5, 10 ** 6, 1, [2, 3]),
])
def test_oracle_sanitized_matches(oracle_san, lo, hi, nearT, psp, bases):
    """The oracle built with the sanitizers gives exactly the plain library's results."""
    import oracle
    rc, out, err = call(oracle_san, lo, hi, nearT, psp, 2, *bases, env=SAN_ENV)
    clean(rc, err)
    o = oracle.scan(lo, hi if hi else 1 << 64, bases=bases, near_T=nearT, psp=bool(psp))
    head = out.splitlines()[0]
    assert head == "primes %d tests %d near %d flt %d nev %d npsp %d" % (o["primes"], o["tests"], o["near"], o["flt"],
                                                                        len(o["events"]), o["npsp"] if psp else 0)
    for b in bases:
        assert "base %d cks %016x sumq %016x" % (b, o["cks"][b], o["sumq"][b]) in out


def test_oracle_sanitized_quotients_top_of_word(oracle_san):
    """The oracle's 128-bit Montgomery (p up to 2^64, carry kept) under the sanitizers: the
    same quotients and residues as the plain library (scanning up there is too slow with
    the sanitizers: the oracle sieves with every prime below 2^32)."""
    import oracle
    ps = [3, 1093, (1 << 63) - 25, (1 << 63) + 29, 18446744073709551557, (1 << 64) - 1, (1 << 62) + 135]
    rc, out, err = call(oracle_san, "q", *ps, env=SAN_ENV)
    clean(rc, err)
    want = []
    for p in ps:
        for b in (2, 3):
            q, r = oracle.fermat_quotient(p, b)
            want.append("%d %d %d %032x" % (p, b, q, r))
    assert out.splitlines() == want
