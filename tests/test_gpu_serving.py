"""Serving (wf_serve_*) on the GPU: continuous batching must give every request exactly what
a scan of the same range gives alone -- totals, checksums and every FOUND / NEAR / FLT
event with the request's own near-hit bound -- and must report, never hide, an event buffer
that overflowed or a step that failed.

The last tests run the serving workload at its full size: 20,000 range queries with Poisson
arrivals at 5,000/s, lengths log-uniform in [1e5, 1e10], starts uniform in [1e9, 1e18].  Every
served result is compared with wf_scan of the same range alone, and a seeded subset of 300
queries of at most 1e7 numbers with the oracle.
"""
import random
import sys
import time

import pytest

import oracle
from mathpub_amd import wf

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def serve(eng, reqs, slots=64, hbm=16 << 30, arrivals=None, limit_s=600):
    """Serve reqs = [(lo, hi, near_T)], submitted at the given arrival times (seconds from the
    start; None = as fast as admission allows).  Returns {index: result or WFError}."""
    eng.serve_start(slots=slots, hbm_budget=hbm)
    got, live, nxt = {}, {}, 0
    t0 = time.time()
    while (nxt < len(reqs) or live) and time.time() - t0 < limit_s:
        now = time.time() - t0
        while nxt < len(reqs) and (arrivals is None or arrivals[nxt] <= now):
            lo, hi, T = reqs[nxt]
            t = eng.serve_submit(lo, hi, near_T=T)
            if t is None:
                break
            live[t] = nxt
            nxt += 1
        for t in list(live):
            try:
                r = eng.serve_poll(t)
            except wf.WFError as ex:
                r = ex
            if r is not None:
                got[live.pop(t)] = r
        time.sleep(0.001)
    eng.serve_stop()
    return got


def same(a, b):
    keys = ("primes", "tests", "near", "flt", "nev", "cks_all", "sumq_all", "events")
    return all(a[k] == b[k] for k in keys)


def test_served_events_equal_scans_alone():
    """Ranges with Wieferich primes (1093, 3511), many near hits, p = 2, the class boundary
    2^50 and the top of the word, each with its own near_T, served together: results and
    events equal wf_scan alone and the oracle."""
    reqs = [(0, 10 ** 4, 0), (0, 10 ** 6, 1000), (1000, 4000, 5), (3500, 3600, 10 ** 6), (2, 3, 0),
            (10 ** 9, 10 ** 9 + 10 ** 6, 10 ** 4), ((1 << 50) - 10 ** 6, (1 << 50) + 10 ** 6, 10 ** 3),
            ((1 << 64) - 10 ** 6, 1 << 64, 10 ** 5), (10 ** 15, 10 ** 15 + 3 * 10 ** 8, 100),
            (10 ** 18, 10 ** 18 + 10 ** 7, 10 ** 9)]
    with wf.Engine() as eng:
        eng.set_bases([2, 5])
        alone = [eng.scan(lo, hi, near_T=T) for lo, hi, T in reqs]
        got = serve(eng, reqs, slots=4, hbm=4 << 30)
    assert len(got) == len(reqs)
    for i, (lo, hi, T) in enumerate(reqs):
        assert same(got[i], alone[i]), reqs[i]
        if hi - lo <= 10 ** 7:
            o = oracle.scan(lo, hi, bases=[2, 5], near_T=T)
            assert (got[i]["primes"], got[i]["near"], got[i]["cks_all"]) == (o["primes"], o["near"], [o["cks"][2], o["cks"][5]])
            assert [(t, p, q, b) for t, p, q, r, b in got[i]["events"]] == [(t, p, q, b) for t, b, p, q, r in o["events"]]
    assert [p for t, p, q, r, b in got[1]["events"] if t == wf.EV_FOUND] == [2, 1093, 3511, 20771, 40487]


def test_served_flt_events():
    """An injected residue fault reaches the served request as its FLT event."""
    lo, hi = 10 ** 15, 10 ** 15 + 10 ** 6
    P = oracle.primes(lo + 1000, lo + 2000)[0]
    with wf.Engine(fault_p=P) as eng:
        want = eng.scan(lo, hi)
        got = serve(eng, [(lo, hi, 0), (lo + 10 ** 6, lo + 2 * 10 ** 6, 0)], slots=2, hbm=2 << 30)
    assert got[0]["flt"] == 1 and same(got[0], want)
    assert [(t, p) for t, p, q, r, b in got[0]["events"]] == [(wf.EV_FLT, P)]
    assert got[1]["flt"] == 0


def test_served_event_overflow_is_reported():
    """A request with more events than a step's event buffer (2^18) fails with
    WF_E_OVERFLOW; the requests served beside it are unaffected."""
    # the other requests have no events, so none of theirs can be among those that did not fit
    reqs = [(10 ** 9, 10 ** 9 + 6 * 10 ** 6, 1 << 62), (10 ** 12, 10 ** 12 + 10 ** 7, 0), (10 ** 15, 10 ** 15 + 10 ** 7, 0)]
    with wf.Engine() as eng:
        alone = [None] + [eng.scan(lo, hi, near_T=T) for lo, hi, T in reqs[1:]]
        got = serve(eng, reqs, slots=4, hbm=4 << 30)
    assert isinstance(got[0], wf.WFError) and got[0].rc == -5
    assert same(got[1], alone[1]) and same(got[2], alone[2])


def test_failed_step_ends_every_request():
    """A serving step that fails (the testing knob fault_step) ends every request with the
    error; new requests are refused, stopping reports it, and the engine scans normally
    afterwards."""
    with wf.Engine(fault_step=2) as eng:
        eng.serve_start(slots=8, hbm_budget=1 << 30)
        tickets = []
        for k in range(4):      # the failure may come before the last submits: they are refused
            try:
                tickets.append(eng.serve_submit(10 ** 12 + k * 10 ** 9, 10 ** 12 + (k + 1) * 10 ** 9))
            except wf.WFError as ex:
                assert ex.rc == -2
        assert tickets
        errs = {}
        t0 = time.time()
        while len(errs) < len(tickets) and time.time() - t0 < 60:
            for t in tickets:
                if t in errs:
                    continue
                try:
                    r = eng.serve_poll(t)
                    if r is not None:
                        errs[t] = 0
                except wf.WFError as ex:
                    errs[t] = ex.rc
            time.sleep(0.005)
        assert sorted(errs.values()) == [-2] * len(tickets)
        with pytest.raises(wf.WFError) as ex:
            eng.serve_submit(10, 100)
        assert ex.value.rc == -2
        with pytest.raises(wf.WFError) as ex:
            eng.serve_stop()
        assert ex.value.rc == -2
        r = eng.scan(10 ** 12, 10 ** 12 + 10 ** 7)
    o = oracle.scan(10 ** 12, 10 ** 12 + 10 ** 7, bases=[2])
    assert (r["primes"], r["cks"]) == (o["primes"], o["cks"][2])


# ---- the serving workload at its full size ----

N_FULL = 20000


def full_queries(n=N_FULL, seed=0x243F6A8885A308D3 ^ 5):
    """n queries: starts uniform in [1e9, 1e18], lengths log-uniform in [1e5, 1e10], and a
    near-hit bound drawn from {0, 100, 10^4, 10^6} per query."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        lo = rng.randrange(10 ** 9, 10 ** 18)
        out.append((lo, lo + int(10 ** rng.uniform(5, 10)), rng.choice([0, 100, 10 ** 4, 10 ** 6])))
    return out


@pytest.fixture(scope="module")
def full_run():
    reqs = full_queries()
    rng = random.Random(17)
    arrivals, t = [], 0.0
    for _ in reqs:
        t += rng.expovariate(5000.0)
        arrivals.append(t)
    eng = wf.Engine()
    eng.scan(10 ** 18, 10 ** 18 + 10 ** 6)          # the sieving primes up to 1e9, once
    t0 = time.time()
    got = serve(eng, reqs, slots=256, hbm=48 << 30, arrivals=arrivals, limit_s=400)
    say("served %d queries at Poisson 5000/s in %.1f s" % (len(got), time.time() - t0))
    yield eng, reqs, got
    eng.close()


@pytest.mark.timeout(600)
def test_full_serving_completes(full_run):
    eng, reqs, got = full_run
    assert len(got) == len(reqs)
    assert not [i for i, r in got.items() if isinstance(r, Exception)]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("part", range(4))
def test_full_serving_equals_scans_alone(full_run, part):
    eng, reqs, got = full_run
    n = len(reqs) // 4
    t0 = time.time()
    for i in range(part * n, (part + 1) * n):
        lo, hi, T = reqs[i]
        a = eng.scan(lo, hi, near_T=T)
        assert same(got[i], a), (i, reqs[i])
    say("part %d: %d queries equal their scans alone (%.1f s)" % (part, n, time.time() - t0))


@pytest.mark.timeout(300)
def test_full_serving_subset_vs_oracle(full_run):
    """300 seeded queries of at most 1e7 numbers against the oracle, run in parallel threads
    (one oracle thread each; its sieving-prime table is built once, for the largest end)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    eng, reqs, got = full_run
    small = [i for i, (lo, hi, T) in enumerate(reqs) if hi - lo <= 10 ** 7]
    pick = random.Random(23).sample(small, 300)
    oracle.scan(max(reqs[i][1] for i in pick) - 1, max(reqs[i][1] for i in pick), bases=[2])
    t0 = time.time()
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        want = list(ex.map(lambda i: oracle.scan(reqs[i][0], reqs[i][1], bases=[2], near_T=reqs[i][2], threads=1), pick))
    say("oracle: %d queries in %.1f s" % (len(pick), time.time() - t0))
    for i, o in zip(pick, want):
        lo, hi, T = reqs[i]
        r = got[i]
        assert (r["primes"], r["tests"], r["near"], r["flt"], r["cks"], r["sumq"]) == \
            (o["primes"], o["tests"], o["near"], o["flt"], o["cks"][2], o["sumq"][2]), reqs[i]
        assert [(t, p, q) for t, p, q, rr, b in r["events"]] == [(e[0], e[2], e[3]) for e in o["events"]]
