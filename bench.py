#!/usr/bin/env python3
"""Benchmark of the MI355X Wieferich engine on the single-range workload.

Workload (one step): every prime p in [1e15, 1e15 + 1e11) sieved on the GPU and
Wieferich-tested for base 2 (q = (2^(p-1) mod p^2 - 1)/p, checksum, sum of q, near list
|A| <= 100); all p are below 2^50, so the FP64 p-adic kernel runs (its class reaches 9 * 2^47).
Inputs are generated on the device (the range itself is the input); nothing crosses
PCIe in the timed region except the ~1 KB of totals and events per chunk.

Prints ONE JSON line:  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LO, SPAN = 10 ** 15, 10 ** 11
NEAR_T = 100
METRIC = "primes Wieferich-tested/sec at p~1e15, batched vs plain GPU, bit-exact cks"
F64_PMAX = 9 << 47                # the engine's FP64 p-adic class is p < 9 * 2^47 (about 1.267e15)
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 vector peak (published), FMA = 2 flops
FP64_PEAK_INSTR = 256 * 64 * 2.4e9   # wave64 FP64 VALU instructions: lanes per second
# 32x32->64 products per second of v_mad_u64_u32 over the whole chip, measured by
# mathpub_amd/wf_probe valu (profiles/valu_rates.txt: 35.47 lane-ops per clock and CU)
INT32_MUL_PEAK = 2.179e13
K_TEST_VALU_PER_TEST = 784           # k_test<A_F64> at [1e15, 1e15+1e11): profiles/c2_sq_counters_pmc.md


def wf_products(pbits):
    """Algorithmic 32x32->64 partial products of one test (the survey's W(p))."""
    L = math.ceil(2 * pbits / 32)
    return (pbits - 1) * (L * (L + 1) // 2 + L + L * L) + (L + L * L)


def fp64_flops(ebits, big=True, W=5):
    """FP64 flops (FMA = 2) and instructions of one base-2 test in the p-adic kernel
    (fp64_pow2x2), exponent length ebits.  A squaring is 2 mul + 7 FMA + 4 add (20 flops, 13
    instructions).  For p >= 2^32 (big): the six leading bits at once (5 flops, 3), then
    ebits - 6 squarings and ceil((ebits - 6) / W) steps x 2^d of 11 flops (9 instructions).
    Below 2^32: the leading W-bit digit, then ceil(ebits / W) - 1 windows of W squarings and
    a 16-flop (12-instruction) step.  Plus the set-up, one division 1/p (12 flops, 10)."""
    if big:
        sq = ebits - 6
        steps = -(-sq // W)
        return sq * 20 + steps * 11 + 5 + 12, sq * 13 + steps * 9 + 3 + 10
    nw = -(-ebits // W)
    sq = (nw - 1) * W
    return sq * 20 + (nw - 1) * 16 + 12, sq * 13 + (nw - 1) * 12 + 10


def cpu_baseline(target_s=8.0):
    """The oracle (a C restatement of wieferich.c's algorithm, 8 lanes, window 5) on all
    usable host cores, on a growing slice of the same workload until it runs >= target_s."""
    import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    span = 10 ** 8
    oracle.scan(LO, LO + 10 ** 6, bases=[2], threads=cores)       # sieving primes, warm caches
    while True:
        t = time.perf_counter()
        r = oracle.scan(LO, LO + span, bases=[2], near_T=NEAR_T, threads=cores)
        dt = time.perf_counter() - t
        if dt >= target_s or span >= SPAN:
            break
        span = min(SPAN, span * max(2, int(target_s / max(dt, 1e-3) * 1.2)))
    return {"value": r["tests"] / dt, "unit": "primes tested/s", "cores": cores, "kind": "port",
            "sample": "[1e15, 1e15+%.3g) base 2, %d primes in %.2f s (oracle/wf_oracle.c)" % (span, r["primes"], dt)}


def c3_ranges(n=4096, seed=0x243F6A8885A308D3):
    """The batched workload: n ranges, starts uniform in [1e12, 1e18], lengths log-uniform in
    [1e6, 1e9], from splitmix64 with a fixed seed (A001220/wieferich.c:1710-1718)."""
    M = (1 << 64) - 1
    st = seed

    def nxt():
        nonlocal st
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    out = []
    for _ in range(n):
        u1 = nxt() / 2.0 ** 64
        u2 = nxt() / 2.0 ** 64
        lo = int(1e12 + u1 * (1e18 - 1e12))
        ln = int(10 ** (6 + 3 * u2))
        out.append((lo, lo + ln))
    return out


def _engine_line(metric, value, unit, ms, args, config, extra=None):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (ranges from a fixed splitmix64 seed; primes sieved on the GPU)", "config": config,
           "roofline": None, "cpu_baseline": None}
    if extra:
        out.update(extra)
    print(json.dumps(out))
    return 0


def int_roofline(kernel, products, ms, launches, note):
    """Roofline of an integer-multiply-bound kernel: algorithmic 32x32 products (the W(p)
    convention) per second of its HIP-event time, against the measured v_mad_u64_u32 rate.
    "mfma" is the compute side of the two-way split: no MFMA runs (one modulus per lane)."""
    achieved = products / (ms / 1e3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": INT32_MUL_PEAK / 1e12, "unit": "Tmul32/s",
            "frac": achieved / (INT32_MUL_PEAK / 1e12), "traffic": None, "kernel": kernel,
            "compute_pipe": "integer VALU (v_mad_u64_u32, measured peak); no MFMA: one modulus per lane",
            "avg_launch_ms": ms / max(launches, 1), "launches": launches, "work": note}


def cpu_sample(pieces, target_s):
    """The oracle (wieferich.c's algorithm) on all usable host cores over the pieces
    [(lo, hi, kwargs)] in order until it has run target_s seconds: primes tested per second
    and what was covered."""
    import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    tests = done = numbers = 0
    el = 0.0
    for lo, hi, kw in pieces:
        t = time.perf_counter()
        r = oracle.scan(lo, hi, threads=cores, **kw)
        el += time.perf_counter() - t
        tests += r["tests"]
        numbers += hi - lo
        done += 1
        if el >= target_s:
            break
    return {"value": tests / el, "unit": "primes tested/s", "cores": cores, "kind": "port",
            "sample": "%d of the workload's ranges in order (%.3g numbers, %d primes tested) in %.1f s "
                      "(oracle/wf_oracle.c)" % (done, numbers, tests, el)}


def _prove(args):
    from mathpub_amd import wf
    return {"auto": wf.PROVE_AUTO, "off": wf.PROVE_OFF, "on": wf.PROVE_ON}[args.prove]


def bench_c3(args):
    """4096 uneven ranges in one batched (HIP graph) call, per-range counts and checksums."""
    from mathpub_amd import wf
    ranges = c3_ranges()
    eng = wf.Engine(arith=args.arith, prove=_prove(args), sieve_block=args.sieve_block, bucket_min=args.bucket_min,
                    seg_bytes=args.seg_bytes)
    t = time.perf_counter()
    tuned = None if args.no_tune else eng.autotune_batch(ranges)
    print("autotune %s (%.1f s)" % (tuned, time.perf_counter() - t), file=sys.stderr, flush=True)
    for i in range(args.warmup):
        t = time.perf_counter()
        per, ev = eng.scan_batch(ranges)
        print("c3 warmup %d: %.2f s" % (i, time.perf_counter() - t), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tims = []
    for _ in range(args.steps):
        per, ev = eng.scan_batch(ranges)
        tims.append(eng.last_timing())
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    tests = sum(r["tests"] for r in per)
    numbers = sum(h - l for l, h in ranges)
    prods = sum(r["tests"] * wf_products(((l + h) // 2).bit_length()) for (l, h), r in zip(ranges, per))
    digest = 0
    for r in per:
        digest = (digest * 0x100000001B3 ^ r["cks"] ^ r["primes"]) & ((1 << 64) - 1)
    # the oracle's per-range counts and checksums of the same 4096 ranges
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "c3_full.json")
    if os.path.exists(gpath) and len(ranges) == 4096:
        with open(gpath) as f:
            g = json.load(f)
        golden = ([r["primes"] for r in per] == g["per_range_primes"]
                  and ["%016x" % r["cks"] for r in per] == g["per_range_cks2"])
    # the plain engine on the same ranges (passes one after another, every kernel timed with
    # HIP events on its stream): the batched/plain ratio and the test kernel's roofline
    plain = None
    roof = None
    if not args.no_plain:
        from mathpub_amd import wf as _wf
        cfg = tuned or {}
        peng = _wf.Engine(arith=args.arith, prove=_prove(args), pipeline=_wf.PIPELINE_OFF,
                          seg_bytes=cfg.get("seg_bytes", args.seg_bytes), sieve_block=cfg.get("sieve_block", args.sieve_block),
                          bucket_min=cfg.get("bucket_min", args.bucket_min))
        t = time.perf_counter()
        pper, _ = peng.scan_batch(ranges)
        pms = (time.perf_counter() - t) * 1e3
        pt = peng.last_timing()
        peng.close()
        assert [(x["primes"], x["cks"]) for x in pper] == [(x["primes"], x["cks"]) for x in per]
        plain = {"ms_per_step": pms, "primes_per_s": tests / (pms / 1e3), "batched_over_plain": pms / ms,
                 "sieve_ms": pt["sieve_ms"], "test_ms": pt["test_ms"], "launches": pt["launches"]}
        roof = int_roofline("k_test<A_M28> (and the one k_test<A_F64> pass of the ranges below 9*2^47)", prods,
                            pt["test_ms"], pt["launches"],
                            "sum over ranges of tests x W(bit length of the range's middle), W = the survey's 32x32 "
                            "product count per test (1730 at 58 bits); HIP-event time of the plain engine's test launches")
        print("c3 plain engine: %s" % plain, file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_sample([(lo, hi, {"bases": [2]}) for lo, hi in ranges], 15.0)
    compare = None
    if args.c3_compare:
        # the same first N ranges through the plain single-range path (one wf_scan each)
        # and through one batched call: the gain of batching many uneven ranges
        sub = ranges[:args.c3_compare]
        t = time.perf_counter()
        single = [eng.scan(lo, hi) for lo, hi in sub]
        t_plain = time.perf_counter() - t
        t = time.perf_counter()
        bper, _ = eng.scan_batch(sub)
        t_batch = time.perf_counter() - t
        assert [(x["primes"], x["cks"]) for x in single] == [(x["primes"], x["cks"]) for x in bper]
        nt = sum(x["tests"] for x in single)
        compare = {"ranges": len(sub), "plain_primes_per_s": nt / t_plain, "batched_primes_per_s": nt / t_batch,
                   "batched_over_plain": t_plain / t_batch}
        print("c3 compare: %s" % compare, file=sys.stderr, flush=True)
    eng.close()
    return _engine_line("primes Wieferich-tested/sec, 4096 batched ranges", tests / (ms / 1e3), "primes tested/s",
                        ms, args, {"workload": "4096 ranges, starts U[1e12,1e18], lengths logU[1e6,1e9], seed "
                                   "0x243F6A8885A308D3", "engine": "batched", "ranges": len(ranges),
                                   "numbers": numbers, "tests": tests, "per_range_digest": "%016x" % digest,
                                   "bit_exact_vs_golden": golden, "autotuned": tuned, "w_products_per_s": prods / (ms / 1e3),
                                   "plain_vs_batched_same_ranges": compare,
                                   "device_ms_per_step": sum(t["total_ms"] for t in tims) / len(tims),
                                   "host_plan_ms_per_step": sum(t["plan_ms"] for t in tims) / len(tims),
                                   "prove": args.prove, "primes_proved_per_step": tims[-1]["proved"],
                                   "plain_engine": plain},
                        {"roofline": roof, "cpu_baseline": cpu})


def bench_c4(args):
    """A slice [2^64 - S, 2^64) of the top-of-word range: full 128-bit Montgomery with carry
    plus the base-2 strong-pseudoprime flags of every odd composite."""
    from mathpub_amd import wf
    span = int(args.c4_span)
    lo = (1 << 64) - span
    eng = wf.Engine(prove=_prove(args))
    for i in range(args.warmup):
        eng.scan(lo, 1 << 64, psp=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = eng.scan(lo, 1 << 64, psp=True)
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    tm = eng.last_timing()
    eng.close()
    # k_psp dominates: W for a base-2 strong test of a 64-bit n is 63 squarings of 9 products
    # (two 32-bit limbs: 3 square, 2 quotient-digit and 4 modulus products) and the final REDC (6)
    w_sprp = 63 * 9 + 6
    roof = int_roofline("k_psp", tm["psp_tests"] * w_sprp, tm["psp_ms"], tm["launches"],
                        "base-2 strong tests run (%d, the numbers the order sieve and the small-factor filter keep) "
                        "x %d products each; HIP-event time of the k_psp launches" % (tm["psp_tests"], w_sprp))
    roof["test_kernel"] = int_roofline("k_test<A_M28>", r["tests"] * wf_products(64), tm["test_ms"], tm["launches"],
                                       "tests x W(64 bits) = %d" % wf_products(64))
    cpu = None
    if not args.no_cpu:
        import oracle
        oracle.scan((1 << 64) - 2, 1 << 64, bases=[2])       # the sieving primes below 2^32, once
        cpu = cpu_sample([((1 << 64) - s, 1 << 64, {"bases": [2], "psp": True}) for s in
                          (10 ** 6, 4 * 10 ** 6, 16 * 10 ** 6, 64 * 10 ** 6, 256 * 10 ** 6)], 10.0)
    return _engine_line("primes Wieferich-tested/sec at p~2^64 with SPRP-2 flags", r["tests"] / (ms / 1e3),
                        "primes tested/s", ms, args,
                        {"workload": "[2^64-%d, 2^64), base 2, strong pseudoprimes to base 2 listed" % span,
                         "engine": "plain", "tests": r["tests"], "cks": "%016x" % r["cks"], "psp": r["psp"],
                         "bpsw_pseudoprimes": r["nbpsw"],
                         "odd_numbers_per_s": span / 2 / (ms / 1e3), "sieve_ms": tm["sieve_ms"],
                         "test_ms": tm["test_ms"], "psp_ms": tm["psp_ms"], "psp_tests": tm["psp_tests"]},
                        {"roofline": roof, "cpu_baseline": cpu})


OEIS_BASES = [3, 5, 6, 7, 10, 12, 13, 14, 15, 17, 18, 19, 20, 22, 23, 26, 30]
# The reference's published speed table (A001220/README.md:116-120): tests/s on an NVIDIA GB10
# for the 16-17 bases at 2e14 and for base 3 at 1.2e15.
PUBLISHED = {"oeis": {"lo": 2 * 10 ** 14, "bases": OEIS_BASES, "gb10": 8.3e8,
                      "metric": "tests/s (prime, base pairs), the 17 'oeis' bases at p~2e14"},
             "base3": {"lo": 12 * 10 ** 14, "bases": [3], "gb10": 7.2e8, "metric": "tests/s, base 3 at p~1.2e15"}}


def bench_published(args):
    """The two workloads of the reference's published speed table, same bases and heights:
    vs_baseline is this engine's rate over the reference GPU tool's (NVIDIA GB10)."""
    from mathpub_amd import wf
    w = PUBLISHED[args.workload]
    lo, span = w["lo"], int(args.span)
    eng = wf.Engine(prove=_prove(args))
    eng.set_bases(w["bases"])
    tuned = None if args.no_tune else eng.autotune(lo, lo + span)
    for _ in range(args.warmup):
        eng.scan_batch([(lo, lo + span)])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per, ev = eng.scan_batch([(lo, lo + span)])
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    r = per[0]
    eng.close()
    value = r["tests"] / (ms / 1e3)
    return _engine_line(w["metric"], value, "tests/s", ms, args,
                        {"workload": "[%.3g, %.3g + %.3g), bases %s" % (lo, lo, span, ",".join(map(str, w["bases"]))),
                         "engine": "batched", "primes_per_step": r["primes"], "tests_per_step": r["tests"],
                         "cks": ["%016x" % c for c in r["cks_all"]], "autotuned": tuned,
                         "published_gb10_tests_per_s": w["gb10"], "published_source": "A001220/README.md:116-120"},
                        {"vs_baseline": value / w["gb10"], "dtype": "f64",
                         "data": "synthetic (the range is the input; primes sieved on the GPU)"})


def bench_prove(args):
    """Prove mode against the full sieve: sieve-stage device ms of one range per call
    (fill + sieve + prove, wf_last_timing) for lengths 1e5..1e9 at heights 1e13..2^64, with
    prove OFF, ON and AUTO; the checksums of the three must agree."""
    from mathpub_amd import wf
    engines = {m: wf.Engine(prove=getattr(wf, "PROVE_" + m.upper())) for m in ("off", "on", "auto")}
    rows, tot = [], {m: 0.0 for m in engines}
    for h in (10 ** 13, 10 ** 15, 10 ** 17, 10 ** 18, (1 << 64) - (1 << 33)):
        for L in (10 ** 5, 10 ** 6, 10 ** 7, 10 ** 8, 10 ** 9):
            row = {"height": h, "length": L}
            cks = set()
            for m, e in engines.items():
                e.scan(h, h + L)
                best = None
                for _ in range(max(args.steps, 1)):
                    r = e.scan(h, h + L)
                    t = e.last_timing()["sieve_ms"]
                    best = t if best is None else min(best, t)
                cks.add((r["primes"], r["cks"]))
                row[m + "_ms"] = round(best, 3)
                tot[m] += best
            assert len(cks) == 1, row
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
    for e in engines.values():
        e.close()
    return _engine_line("sieve-stage time of one range per call, full sieve over auto prove mode", tot["off"] / tot["auto"],
                        "x", tot["auto"], args,
                        {"workload": "one range per call, lengths 1e5..1e9 at heights 1e13..2^64-2^33",
                         "total_ms": {m: round(v, 3) for m, v in tot.items()}, "rows": rows},
                        {"dtype": "u64", "data": "synthetic (the ranges are the input)"})


def bench_stats(args):
    """wf_sieve_stats over [0, 1e12): twin pairs, largest gap, A158939 run lengths (the sieve
    exposed to the other prime consumers), checked against the reference's data."""
    from mathpub_amd import wf
    hi = 10 ** 12
    eng = wf.Engine()
    for _ in range(args.warmup):
        eng.sieve_stats(0, 10 ** 10)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = eng.sieve_stats(0, hi)
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    eng.close()
    ok = st["twins"] == 1870585220 and st["primes"] == 37607912018 and st["first"].get(13) == 968068681511
    return _engine_line("numbers sieved/s with twin, gap and run statistics", hi / (ms / 1e3), "numbers/s", ms, args,
                        {"workload": "[0, 1e12)", "primes": st["primes"], "twins": st["twins"],
                         "max_gap": st["max_gap"], "max_gap_p": st["max_gap_p"],
                         "first": {str(k): v for k, v in st["first"].items()},
                         "matches_reference_data": ok},
                        {"dtype": "u32", "data": "synthetic (the range is the input)"})


def c5_queries(n, seed=0x243F6A8885A308D3 ^ 5):
    """n range queries: lengths log-uniform in [1e5, 1e10], starts uniform in [1e9, 1e18]."""
    import random
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        lo = rng.randrange(10 ** 9, 10 ** 18)
        out.append((lo, lo + int(10 ** rng.uniform(5, 10))))
    return out


def bench_c5(args):
    """Serving: Poisson arrivals at each rate, continuous batching in the library, latency
    from arrival to completion (admission waits included)."""
    import random
    from mathpub_amd import wf
    eng = wf.Engine(prove=_prove(args))
    eng.scan(10 ** 18, 10 ** 18 + 10 ** 6)          # device-resident sieving primes up to 1e9
    qs = c5_queries(args.queries)
    rng = random.Random(17)
    sweep = []
    for rate in [float(x) for x in args.rates.split(",")]:
        eng.serve_start(slots=256, hbm_budget=48 << 30)
        arrivals, t = [], 0.0
        for _ in qs:
            t += rng.expovariate(rate)
            arrivals.append(t)
        lat, live, nxt, done_tests, numbers, prods = [], {}, 0, 0, 0, 0
        t0 = time.perf_counter()
        while nxt < len(qs) or live:
            now = time.perf_counter() - t0
            while nxt < len(qs) and arrivals[nxt] <= now:
                tk = eng.serve_submit(qs[nxt][0], qs[nxt][1], near_T=NEAR_T)
                if tk is None:
                    break                                  # full: the request waits (its latency grows)
                live[tk] = nxt
                nxt += 1
            for tk in list(live):
                r = eng.serve_poll(tk)
                if r is not None:
                    i = live.pop(tk)
                    lat.append(time.perf_counter() - t0 - arrivals[i])
                    done_tests += r["tests"]
                    numbers += qs[i][1] - qs[i][0]
                    prods += r["tests"] * wf_products(((qs[i][0] + qs[i][1]) // 2).bit_length())
                    if len(lat) % 2000 == 0:
                        print("rate %g: %d queries done, %.1f s" % (rate, len(lat), time.perf_counter() - t0),
                              file=sys.stderr, flush=True)
            time.sleep(0.0005)
        wall = time.perf_counter() - t0
        eng.serve_stop()
        tm = eng.last_timing()
        lat.sort()
        sweep.append({"rate_per_s": rate, "queries": len(lat), "p50_s": lat[len(lat) // 2],
                      "p99_s": lat[min(len(lat) - 1, int(len(lat) * 0.99))], "wall_s": wall,
                      "primes_tested_per_s": done_tests / wall, "numbers_per_s": numbers / wall,
                      "test_ms": tm["test_ms"], "sieve_ms": tm["sieve_ms"], "test_launches": tm["launches"],
                      "products": prods})
        print("rate %g: %s" % (rate, sweep[-1]), file=sys.stderr, flush=True)
    eng.close()
    best = max(x["primes_tested_per_s"] for x in sweep)
    top = max(sweep, key=lambda x: x["primes_tested_per_s"])
    roof = int_roofline("k_test<A_M28> (serving steps)", top["products"], top["test_ms"], top["test_launches"],
                        "at %g queries/s: sum over queries of tests x W(bit length of the query's middle); HIP-event "
                        "time of the steps' test launches" % top["rate_per_s"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_sample([(lo, hi, {"bases": [2], "near_T": NEAR_T}) for lo, hi in qs], 15.0)
    return _engine_line("serving: primes Wieferich-tested/sec and p50/p99 latency", best, "primes tested/s",
                        sweep[-1]["wall_s"] * 1e3, args,
                        {"workload": "%d queries per rate, lengths logU[1e5,1e10], starts U[1e9,1e18], Poisson "
                                     "arrivals" % len(qs), "engine": "serving (continuous batching)",
                         "prove": args.prove, "sweep": sweep},
                        {"roofline": roof, "cpu_baseline": cpu})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--arith", type=int, default=0, help="1 = force the 32-bit Montgomery kernels")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="c3: skip the plain-engine run (and the roofline)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (engine defaults)")
    ap.add_argument("--seg-bytes", type=int, default=0, help="c2/c3 with --no-tune: LDS segment bytes")
    ap.add_argument("--sieve-block", type=int, default=0, help="c2/c3 with --no-tune: threads per sieve workgroup")
    ap.add_argument("--bucket-min", type=int, default=0, help="c2/c3 with --no-tune: bucket threshold")
    ap.add_argument("--span", type=float, default=SPAN, help="numbers per step (default 1e11)")
    ap.add_argument("--mode", choices=["both", "plain", "batched"], default="both")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5", "oeis", "base3", "prove", "stats"], default="c2",
                    help="c2 = the single range (default, the headline line); c3 = 4096 batched ranges; "
                         "c4 = a slice of the top-of-word range with strong-pseudoprime flags; c5 = serving; "
                         "oeis / base3 = the reference's published speed workloads (17 bases at 2e14, base 3 at 1.2e15)")
    ap.add_argument("--queries", type=int, default=2000, help="c5: queries per rate")
    ap.add_argument("--rates", default="100,1000,5000", help="c5: Poisson arrival rates per second")
    ap.add_argument("--c4-span", type=float, default=2.0 ** 34, help="c4: numbers below 2^64 per step")
    ap.add_argument("--c3-compare", type=int, default=0,
                    help="c3: also time the first N ranges one wf_scan at a time against one batched call")
    ap.add_argument("--prove", choices=["auto", "off", "on"], default="auto",
                    help="prove mode: short pieces sieved with small primes, survivors decided by BPSW")
    ap.add_argument("--test-wgs", type=int, default=0,
                    help="batched: test workgroups per CU (0 = engine default)")
    args = ap.parse_args()
    if args.workload != "c2" and args.gpus == 1:
        return {"c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "oeis": bench_published,
                "base3": bench_published, "prove": bench_prove, "stats": bench_stats}[args.workload](args)
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this build "
              "(the range would shard into contiguous sub-ranges, one engine per GPU, with the "
              "totals summed mod 2^64 at the end)", file=sys.stderr)
        return 2
    from mathpub_amd import wf
    span = int(args.span)
    eng = wf.Engine(arith=args.arith, test_wgs=args.test_wgs, prove=_prove(args), sieve_block=args.sieve_block,
                    bucket_min=args.bucket_min, seg_bytes=args.seg_bytes)
    print("engine: " + eng.info(), file=sys.stderr, flush=True)

    def timed(fn, label):
        ref = None
        for i in range(args.warmup):
            t = time.perf_counter()
            ref = fn()
            print("%s warmup %d: %.3f s, %d primes, cks %016x" % (label, i, time.perf_counter() - t, ref["primes"],
                                                                  ref["cks"]), file=sys.stderr, flush=True)
        res, tim = [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            res.append(fn())
            tim.append(eng.last_timing())
        el = time.perf_counter() - t0
        print("%s: %.3f s per step" % (label, el / args.steps), file=sys.stderr, flush=True)
        return res + ([ref] if ref else []), tim, el / args.steps * 1e3

    plain = batched = None
    tuned = None
    if not args.no_tune:
        # both engines run the tuned configuration, so that their ratio is the pipeline's alone
        t = time.perf_counter()
        tuned = eng.autotune(LO, LO + span)
        print("autotune: %s (%.1f s)" % (tuned, time.perf_counter() - t), file=sys.stderr, flush=True)
    if args.mode in ("both", "plain"):
        plain = timed(lambda: eng.scan(LO, LO + span, near_T=NEAR_T), "plain")
    if args.mode in ("both", "batched"):

        def one():
            per, ev = eng.scan_batch([(LO, LO + span)], near_T=NEAR_T)
            return per[0]
        batched = timed(one, "batched")
    main_run = batched or plain
    results, timings, ms_per_step = main_run
    r = results[0]
    same = all((x["primes"], x["cks"], x["sumq"], x["near"]) == (r["primes"], r["cks"], r["sumq"], r["near"])
               for x in results + (plain[0] if plain else []))
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "c2_full.json")
    if span == SPAN and os.path.exists(gpath):
        g = json.load(open(gpath))
        golden = (g["primes"] == r["primes"] and g["cks2"] == "%016x" % r["cks"]
                  and g["sumq2"] == "%016x" % r["sumq"] and g["near"] == r["near"])
    tests = r["tests"]
    value = tests / (ms_per_step / 1e3)
    roof = None
    if plain:
        # dominant kernel: the Fermat-quotient test, timed with HIP events on the engine's
        # stream in the plain path (one sieve + one test launch per pass)
        ptim = plain[1]
        test_ms = sum(t["test_ms"] for t in ptim) / args.steps
        sieve_ms = sum(t["sieve_ms"] for t in ptim) / args.steps
        launches = max(1, ptim[-1]["launches"])
        pbits = (LO + span // 2).bit_length()
        flops, instr = fp64_flops(pbits, big=LO >= (1 << 32))
        if args.arith == 1 or (LO + span) > F64_PMAX:
            achieved = tests * wf_products(pbits) / (test_ms / 1e3) / 1e12
            # compute-bound on the integer multiplier: the peak is the measured issue rate of
            # v_mad_u64_u32 (no published figure exists for it)
            roof = {"bound": "mfma", "achieved": achieved, "peak": INT32_MUL_PEAK / 1e12, "unit": "Tmul32/s",
                    "frac": achieved / (INT32_MUL_PEAK / 1e12), "traffic": None, "kernel": "k_test<A_M4>",
                    "compute_pipe": "integer VALU (v_mad_u64_u32); no MFMA: one modulus per lane"}
        else:
            achieved = tests * flops / (test_ms / 1e3) / 1e12
            # compute-bound ("mfma" in the roofline's two-way split): the FP64 peak, 78.6 TFLOP/s,
            # is MI355X's published FP64 figure; the kernel issues it on the vector pipe, since
            # every lane has its own modulus and there is no matrix product
            roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_test<A_F64>",
                    "compute_pipe": "FP64 VALU (v_fma_f64); no MFMA: one modulus per lane",
                    "valu_fp64_issue_frac": tests * instr / (test_ms / 1e3) / FP64_PEAK_INSTR,
                    # all VALU instructions of the kernel (SQ_INSTS_VALU x 64 / tests, counted once
                    # with rocprofv3 --pmc on this workload) against the same nominal-clock ceiling
                    "valu_instr_per_test": K_TEST_VALU_PER_TEST,
                    "valu_issue_frac": tests * K_TEST_VALU_PER_TEST / (test_ms / 1e3) / FP64_PEAK_INSTR}
        tj = os.path.join(ROOT, "profiles", "c2_k_test_traffic.json")
        if roof["kernel"] == "k_test<A_F64>" and span == SPAN and os.path.exists(tj):
            # HBM bytes per launch from the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
            # this workload, scaled by the same counters' reading of a plain 1 GiB copy
            roof["traffic"] = json.load(open(tj))["hbm_bytes_per_launch"]
            roof["traffic_unit"] = "bytes per launch (rocprofv3 --pmc, profiles/c2_hbm_traffic_pmc.md)"
        roof["avg_launch_ms"] = test_ms / launches
        roof["launches_per_step"] = launches
        roof["w_products_per_s"] = tests * wf_products(pbits) / (test_ms / 1e3)
        roof["sieve_ms_per_step"] = sieve_ms
        roof["test_ms_per_step"] = test_ms
        roof["host_plan_ms_per_step"] = sum(t["plan_ms"] for t in ptim) / args.steps
    cfg = {"workload": "single range [1e15, 1e15+%.3g), base 2, near |A|<=%d" % (span, NEAR_T),
           "engine": "batched" if batched else "plain", "primes_per_step": r["primes"], "tests_per_step": tests,
           "cks": "%016x" % r["cks"], "sumq": "%016x" % r["sumq"], "near": r["near"],
           "bit_exact_across_steps_and_engines": same, "bit_exact_vs_golden": golden, "parallelism": "1 GPU",
           "autotuned": tuned}
    if plain and batched:
        pv = tests / (plain[2] / 1e3)
        cfg["plain_value"] = pv
        cfg["plain_ms_per_step"] = plain[2]
        cfg["batched_over_plain"] = value / pv
    if batched:
        cfg["batched_host_plan_ms_per_step"] = sum(t["plan_ms"] for t in batched[1]) / args.steps
        # the graph's own time (HIP events around the launch); the rest of ms_per_step is host
        # work: planning, descriptor copies, graph capture and instantiation, results
        cfg["batched_device_ms_per_step"] = sum(t["total_ms"] for t in batched[1]) / args.steps
    out = {
        "metric": METRIC, "value": value, "unit": "primes tested/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if (roof or {}).get("unit") == "TFLOP/s" else "u32",
        "data": "synthetic (the range is the input; primes sieved on the GPU)",
        "config": cfg, "roofline": roof, "cpu_baseline": None,
    }
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
