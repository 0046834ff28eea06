"""The wf host program: command line, number grammar, checkpoint validation and output
formats of A001220/wieferich.c (CPU tests need no GPU: they stop before the engine opens;
the gpu-marked tests scan for real)."""
import os
import re
import signal
import subprocess
import sys
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WF = os.path.join(ROOT, "mathpub_amd", "wf")
pytestmark_cpu = pytest.mark.skipif(not os.path.exists(WF), reason="wf not built")


def run(*args, timeout=600):
    p = subprocess.run([WF] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr


@pytestmark_cpu
def test_usage_exit_2():
    assert run()[0] == 2
    assert run("frobnicate")[0] == 2
    assert run("scan")[0] == 2
    assert run("scan", "1", "2", "3")[0] == 2
    assert run("scan", "10", "-z")[0] == 2


@pytestmark_cpu
@pytest.mark.parametrize("arg,msg", [("1e-5", "bad number"), ("e5", "bad number"), ("1.5e0", "bad number"),
                                     ("2^65", "bad number"), ("5-7", "negative result"), ("1-2-3", "bad number"),
                                     ("18446744073709551616+1", "overflow")])
def test_number_grammar_errors(arg, msg):
    rc, out, err = run("scan", arg, "2^70")
    rc2, out2, err2 = run("scan", "0", arg)
    assert rc == 1 or rc2 == 1
    assert msg in err + err2


@pytestmark_cpu
def test_range_errors():
    assert "END must be larger than START" in run("scan", "10", "5")[2]
    assert "START must be < 2^64" in run("scan", "2^64", "2^64")[2]
    rc, out, err = run("scan", "2^64-10", "2^64+0", "-c", "7")
    assert rc == 1 and "END must be <= 2^64" in err          # rounded up past 2^64
    assert "P must be odd" in run("check", "4")[2]
    assert "bad base" in run("scan", "10", "-b", "1")[2]


def write_ckpt(path, **kw):
    f = {"bases": "2", "start": 1000, "chunk": 100, "end": 2000, "frontier": 1, "primes": 5, "tests": 5,
         "near": 0, "flt": 0}
    f.update(kw)
    lines = ["wieferich checkpoint 1"]
    for k in ("bases", "start", "chunk", "end", "frontier", "primes", "tests", "near", "flt"):
        if f[k] is not None:
            lines.append("%s %s" % (k, f[k]))
    if kw.get("cks", True):
        lines.append("cks 2 0123456789abcdef")
    if kw.get("terminated", True):
        lines.append("end")
    open(path, "w").write("\n".join(lines) + "\n")


@pytestmark_cpu
@pytest.mark.parametrize("kw,msg", [({"terminated": False}, "is truncated"), ({"cks": False}, "has no checksum"),
                                    ({"start": 999}, "was written for start=999"),
                                    ({"chunk": 50}, "chunk=50"), ({"bases": "3"}, "was written for bases 3"),
                                    ({"frontier": 99}, "is past the new END")])
def test_checkpoint_validation(tmp_path, kw, msg):
    ck = tmp_path / "st"
    write_ckpt(ck, **kw)
    rc, out, err = run("scan", "1000", "2000", "-c", "100", "-b", "2", "-S", ck)
    assert rc == 1 and msg in err, err


@pytestmark_cpu
def test_foreign_file_rejected(tmp_path):
    ck = tmp_path / "st"
    ck.write_text("hello\n")
    rc, out, err = run("scan", "1000", "2000", "-c", "100", "-S", ck)
    assert rc == 1 and "not a wieferich checkpoint" in err


# ---- on the GPU ----

def parse_summary(out):
    """SUMMARY totals and the per-base checksum lines ((base, cks), ...)."""
    m = re.search(r"SUMMARY \[(\d+), (\d+)\) primes (\d+) tests (\d+) flt_errors (\d+)", out)
    cks = tuple(re.findall(r"^base +(\d+) +(?:A\d+ +)?cks ([0-9a-f]{16})", out, re.M))
    return m.groups(), cks


@pytest.mark.gpu
def test_scan_matches_oracle_and_log_format(tmp_path):
    import oracle
    log = tmp_path / "log"
    rc, out, err = run("scan", "1e15", "1e15+3e8", "-c", "1e8", "-b", "2", "-n", "1000", "-L", log, "-q", "-Q")
    assert rc == 0, err
    (lo, hi, primes, tests, flt), cks = parse_summary(out)
    o = oracle.scan(10 ** 15, 10 ** 15 + 3 * 10 ** 8, bases=[2], near_T=1000)
    assert int(primes) == o["primes"] and cks == (("2", "%016x" % o["cks"][2]),) and int(flt) == 0
    assert "sumq base 2 %016x" % o["sumq"][2] in out
    lines = log.read_text().splitlines()
    assert len(lines) == 3
    tot = 0
    for k, ln in enumerate(lines):
        a = 10 ** 15 + k * 10 ** 8
        oc = oracle.scan(a, a + 10 ** 8, bases=[2])
        assert ln == "C %d %d %d 2:%016x" % (a, a + 10 ** 8, oc["primes"], oc["cks"][2])
        tot += oc["primes"]
    assert tot == o["primes"]
    near = [l for l in out.splitlines() if l.startswith("NEAR")]
    assert len(near) == o["near"]


@pytest.mark.gpu
def test_found_and_confirmed_lines():
    rc, out, err = run("scan", "0", "1e5", "-c", "2e4", "-b", "2", "-q")
    assert rc == 0
    assert "FOUND base 2 p 1093 (A001220)" in out and "FOUND base 2 p 3511 (A001220)" in out
    assert "CONFIRMED base 2 p 1093 (A001220) (all primes in [0, 20000) tested)" in out
    assert re.search(r"base   2 A001220  cks [0-9a-f]{16}  solutions: 1093 3511", out)


@pytest.mark.gpu
def test_resume_equals_straight_run(tmp_path):
    ck = tmp_path / "st"
    straight = run("scan", "1e12", "1e12+4e8", "-c", "5e7", "-q")[1]     # default bases: oeis
    r1 = run("scan", "1e12", "1e12+2e8", "-c", "5e7", "-q", "-S", ck)
    assert r1[0] == 0 and ck.exists()
    r2 = run("scan", "1e12", "1e12+4e8", "-c", "5e7", "-q", "-S", ck)      # END may change on resume
    assert r2[0] == 0
    assert parse_summary(r2[1]) == parse_summary(straight)


@pytest.mark.gpu
def test_sigint_saves_checkpoint_then_resumes(tmp_path):
    ck = tmp_path / "st"
    straight = run("scan", "1e15", "1e15+3e10", "-c", "1e9", "-b", "2", "-q")[1]
    p = subprocess.Popen([WF, "scan", "1e15", "1e15+3e10", "-c", "1e9", "-b", "2", "-S", str(ck)], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    time.sleep(3.0)
    p.send_signal(signal.SIGINT)
    out, err = p.communicate(timeout=300)
    assert p.returncode == 0 and ck.exists()
    assert "stopped at position" in err or "SUMMARY [1000000000000000, 1000030000000000)" in out
    r2 = run("scan", "1e15", "1e15+3e10", "-c", "1e9", "-b", "2", "-q", "-S", ck)
    assert parse_summary(r2[1]) == parse_summary(straight)


@pytest.mark.gpu
def test_check_command():
    rc, out, err = run("check", "1093")
    assert rc == 0 and "q = 0  A = 0  <-- WIEFERICH  [reference agrees]" in out
    rc, out, err = run("check", "18446744073709551557")
    assert rc == 0 and "(prime)" in out and "agrees" in out
    rc, out, err = run("check", "2047")
    assert rc == 0 and "NOT prime" in out


@pytest.mark.gpu
def test_selftest_small():
    rc, out, err = run("selftest", "1e8", timeout=900)
    assert rc == 0 and "selftest passed" in err, err


OEIS = [3, 5, 6, 7, 10, 12, 13, 14, 15, 17, 18, 19, 20, 22, 23, 26, 30]


@pytest.mark.gpu
def test_default_oeis_bases_match_oracle(tmp_path):
    """The reference's default base list ("oeis", 17 bases): SUMMARY checksums and every
    chunk-log line agree with the oracle, byte for byte."""
    import oracle
    log = tmp_path / "log"
    rc, out, err = run("scan", "1e12", "1e12+2e8", "-c", "1e8", "-n", "100", "-L", log, "-q")
    assert rc == 0, err
    (lo, hi, primes, tests, flt), cks = parse_summary(out)
    o = oracle.scan(10 ** 12, 10 ** 12 + 2 * 10 ** 8, bases=OEIS, near_T=100)
    assert int(primes) == o["primes"] and int(tests) == o["tests"] and int(flt) == 0
    assert cks == tuple((str(b), "%016x" % o["cks"][b]) for b in OEIS)
    for k, ln in enumerate(log.read_text().splitlines()):
        a = 10 ** 12 + k * 10 ** 8
        oc = oracle.scan(a, a + 10 ** 8, bases=OEIS)
        assert ln == "C %d %d %d " % (a, a + 10 ** 8, oc["primes"]) + " ".join(
            "%d:%016x" % (b, oc["cks"][b]) for b in OEIS)
    near = [l for l in out.splitlines() if l.startswith("NEAR")]
    assert len(near) == o["near"]


@pytest.mark.gpu
def test_check_known_bases():
    rc, out, err = run("check", "1006003")
    assert rc == 0 and "base   3: q = 0  A = 0  <-- WIEFERICH  [reference agrees]" in out, out
    assert "DISAGREES" not in out


@pytest.mark.gpu
def test_num_command():
    """wf num: the GPU's base-2 strong test, strong Lucas test and verdict per number."""
    rc, out, err = run("num", "2047", "5459", "1000000007", "18446744073709551557", "18446744073709551615")
    assert rc == 0, err
    assert out.splitlines() == ["2047: gpu sprp2 1 slprp 0 prime 0", "5459: gpu sprp2 0 slprp 1 prime 0",
                                "1000000007: gpu sprp2 1 slprp 1 prime 1",
                                "18446744073709551557: gpu sprp2 1 slprp 1 prime 1",
                                "18446744073709551615: gpu sprp2 0 slprp 0 prime 0"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["on", "off"])
def test_scan_prove_modes_same_output(tmp_path, mode):
    """--prove on / off: the same SUMMARY and chunk log as the default."""
    outs = []
    for m in ("auto", mode):
        log = tmp_path / ("log_" + m)
        rc, out, err = run("scan", "1e17", "1e17+4e6", "-c", "1e6", "-n", "1000", "-L", log, "-q", "-Q", "--prove", m)
        assert rc == 0, err
        outs.append((out, log.read_text()))
    assert outs[0] == outs[1]


@pytestmark_cpu
def test_dcheck_usage():
    rc, out, err = run("dcheck")
    assert rc == 2 and "usage" in err
    rc, out, err = run("dcheck", "/nonexistent/log", "2", "0")
    assert rc == 1 and "STEP must be positive" in err


@pytest.mark.gpu
def test_dcheck_matches_then_flags_a_bad_line(tmp_path):
    """wf dcheck (dcheck.sh's double check on the GPU, with prove mode, 16 KiB segments and the
    32-bit ladders): every chunk of a scan's log matches; a line with one wrong checksum digit
    is reported as MISMATCH; a second run over the same OUT checks nothing again."""
    log = tmp_path / "log"
    rc, out, err = run("scan", "1e15", "1e15+6e7", "-c", "1e7", "-b", "2", "-L", log, "-q")
    assert rc == 0, err
    lines = log.read_text().splitlines()
    assert len(lines) == 6
    out1 = tmp_path / "out1"
    rc, out, err = run("dcheck", log, "2", "1", "8", out1)
    assert rc == 0, err
    res = out1.read_text().splitlines()
    assert [r.split()[:3] for r in res] == [["MATCH"] + l.split()[1:3] for l in lines]
    rc, out, err = run("dcheck", log, "2", "1", "8", out1)       # resumed: nothing left to check
    assert rc == 0 and "0 chunks checked" in err and len(out1.read_text().splitlines()) == 6
    bad = list(lines)
    last = bad[3][-1]
    bad[3] = bad[3][:-1] + ("0" if last != "0" else "1")
    badlog = tmp_path / "badlog"
    badlog.write_text("\n".join(bad) + "\n")
    out2 = tmp_path / "out2"
    rc, out, err = run("dcheck", badlog, "2", "2", "8", out2)     # every 2nd chunk: lines 2, 4, 6
    assert rc == 1
    res = out2.read_text().splitlines()
    assert [r.split()[0] for r in res] == ["MATCH", "MISMATCH", "MATCH"]
    assert res[1].startswith("MISMATCH %s %s gpu: %s check: %s" % (bad[3].split()[1], bad[3].split()[2], bad[3],
                                                                    lines[3]))


@pytest.mark.gpu
def test_dcheck_other_bases_and_top_of_word(tmp_path):
    """The 'oeis' base list at 1e12 (the 32-bit binary ladder against the scan's FP64 windows),
    prime bases across 2^63 (against the radix-2^28 windows), and base 2 just below 2^64
    (END = 2^64 in the log)."""
    for start, end, chunk, bases in (("1e12", "1e12+4e6", "2e6", "oeis"), ("2^63-2e6", "2^63+2e6", "2e6", "3,5,7,11,13"),
                                     ("2^64-4e6", "2^64", "2e6", "2")):
        log = tmp_path / ("log" + bases)
        rc, out, err = run("scan", start, end, "-c", chunk, "-b", bases, "-L", log, "-q")
        assert rc == 0, err
        outp = tmp_path / ("out" + bases)
        rc, out, err = run("dcheck", log, bases, "1", "1", outp)
        assert rc == 0, err
        assert [r.split()[0] for r in outp.read_text().splitlines()] == ["MATCH", "MATCH"]


# ---- every output line pinned, the FLT-error path, event overflow ----

M64 = (1 << 64) - 1
ANUM = {2: "A001220", 3: "A014127", 5: "A123692"}


def expected_stdout(start, end, chunk, bases, near_T, oracle):
    """The scan's whole stdout built from the oracle's results in the reference's formats:
    per chunk its FOUND and NEAR lines in (p, base) order (wieferich.c:1009, :1071), then a
    CONFIRMED line per solution of the chunk (:1350-1352); then SUMMARY (:1505-1523)."""
    out, hits, tot = [], [], {"primes": 0, "tests": 0, "near": 0, "cks": {b: 0 for b in bases}}
    for lo in range(start, end, chunk):
        hi = min(lo + chunk, end)
        o = oracle.scan(lo, hi, bases=bases, near_T=near_T)
        found = []
        for t, b, p, q, r in o["events"]:
            if t == 0:
                out.append("FOUND base %d p %d (%s)" % (b, p, ANUM[b]))
                found.append((b, p))
            elif t == 1:
                out.append("NEAR base %d p %d A %d" % (b, p, q if q <= p // 2 else q - p))
        for b, p in found:
            out.append("CONFIRMED base %d p %d (%s) (all primes in [%d, %d) tested)" % (b, p, ANUM[b], start, hi))
        hits += found
        for k in ("primes", "tests", "near"):
            tot[k] += o[k]
        for b in bases:
            tot["cks"][b] = (tot["cks"][b] + o["cks"][b]) & M64
    out.append("SUMMARY [%d, %d) primes %d tests %d flt_errors 0" % (start, end, tot["primes"], tot["tests"]) +
               (" near(|A|<=%d) %d" % (near_T, tot["near"]) if near_T > 0 else ""))
    for b in bases:
        sol = [str(p) for bb, p in hits if bb == b]
        out.append("base %3d %-8s cks %016x  solutions: %s" % (b, ANUM[b], tot["cks"][b], " ".join(sol) or "none"))
    return "\n".join(out) + "\n", tot, hits


@pytest.mark.gpu
def test_whole_output_and_checkpoint_text(tmp_path):
    """wf scan's complete stdout and the complete checkpoint it leaves, byte for byte."""
    import oracle
    ck = tmp_path / "st"
    rc, out, err = run("scan", "0", "2e5", "-c", "5e4", "-b", "2,3,5", "-n", "3", "-q", "-S", ck)
    assert rc == 0, err
    want, tot, hits = expected_stdout(0, 200000, 50000, [2, 3, 5], 3, oracle)
    assert out == want
    text = ("wieferich checkpoint 1\nbases 2,3,5\nstart 0\nchunk 50000\nend 200000\nfrontier 4\nprimes %d\n"
            "tests %d\nnear %d\nflt 0\n" % (tot["primes"], tot["tests"], tot["near"]) +
            "".join("cks %d %016x\n" % (b, tot["cks"][b]) for b in (2, 3, 5)))
    text += "".join("hit %d %d %d\n" % (b, p, p // 50000) for b, p in hits) + "end\n"
    assert ck.read_text() == text
    # resuming a complete scan re-announces its solutions and prints the same SUMMARY
    rc, out2, err = run("scan", "0", "2e5", "-c", "5e4", "-b", "2,3,5", "-n", "3", "-q", "-S", ck)
    assert rc == 0 and out2.splitlines()[-4:] == out.splitlines()[-4:]
    assert [l for l in out2.splitlines() if l.startswith("CONFIRMED")] == \
        ["CONFIRMED base %d p %d (%s) (all primes in [0, 200000) tested)" % (b, p, ANUM[b]) for b, p in hits]


@pytest.mark.gpu
@pytest.mark.parametrize("lo", [10 ** 15, 10 ** 18])
def test_flt_error_line_and_exit_status(tmp_path, lo):
    """--fault-p P corrupts P's residue on the GPU (r -> r + 1): the scan prints the reference's
    FLT-ERROR line with that residue, counts the pair in tests and flt_errors, leaves it out of
    the checksum, the sum of q and the chunk log, and exits with status 3
    (wieferich.c:1058-1063, :1591)."""
    import oracle
    hi = lo + 2 * 10 ** 6
    P = oracle.primes(lo + 1500000, lo + 1510000)[2]
    r = pow(2, P - 1, P * P)
    q = r // P
    log = tmp_path / "log"
    rc, out, err = run("scan", lo, hi, "-c", "1e6", "-b", "2", "-q", "-Q", "-L", log, "--fault-p", P)
    assert rc == 3, err
    assert "FLT-ERROR base 2 p %d r = %032x" % (P, r + 1) in out.splitlines()
    o = oracle.scan(lo, hi, bases=[2])
    cks = (o["cks"][2] - oracle.mix64(q ^ (P * 0x9E3779B97F4A7C15 & M64))) & M64
    assert "SUMMARY [%d, %d) primes %d tests %d flt_errors 1" % (lo, hi, o["primes"], o["tests"]) in out
    assert re.search(r"^base   2 A001220  cks %016x  solutions: none$" % cks, out, re.M)
    assert "sumq base 2 %016x" % ((o["sumq"][2] - q) & M64) in out
    term = oracle.mix64(q ^ (P * 0x9E3779B97F4A7C15 & M64))      # P's checksum term, left out
    first, second = log.read_text().splitlines()
    o2 = oracle.scan(lo + 10 ** 6, hi, bases=[2])
    assert second == "C %d %d %d 2:%016x" % (lo + 10 ** 6, hi, o2["primes"], (o2["cks"][2] - term) & M64)


@pytest.mark.gpu
def test_event_overflow_message():
    """More events in one chunk than the host's buffer of 65536 (every prime a NEAR event
    under -n 1e18) ends the scan with a message and status 1, never a silent truncation."""
    rc, out, err = run("scan", "1e9", "1e9+2e6", "-c", "2e6", "-b", "2", "-n", "1e18", "-q")
    assert rc == 1 and "events exceed the event buffer (65536); use a smaller -n or CHUNK" in err, err
