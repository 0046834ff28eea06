"""ctypes mirror of include/wf.h.

Names follow the reference's domain: a scan of [lo, hi) returns the totals of
A001220/wieferich.c's slot_t (primes, tests, near, flt, per-base checksums) plus the
sum of the quotients, and FOUND / NEAR / FLT-ERROR events sorted by (p, base).
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(PKG, "libwf.so")
MAXB = 64
EV_FOUND, EV_NEAR, EV_FLT = 0, 1, 2
PROVE_AUTO, PROVE_OFF, PROVE_ON = 0, 1, 2
PIPELINE_GRAPH, PIPELINE_OFF = 0, 1
DERIVE_AUTO, DERIVE_OFF = 0, 1     # composite bases from their prime factors' quotients
BPSW_SPRP2, BPSW_LUCAS, BPSW_PRIME = 1, 2, 4
STATIC_QUOT_MAX = 1 << 18          # wf_quotients_static's fixed scratch area (values per call)

u64, u32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64


class Event(ctypes.Structure):
    _fields_ = [("type", u32), ("j", u32), ("p", u64), ("q", u64), ("r_hi", u64), ("r_lo", u64), ("range", u32),
                ("pad", u32)]


class Result(ctypes.Structure):
    _fields_ = [("primes", u64), ("tests", u64), ("near", u64), ("flt", u64), ("cks", u64 * MAXB),
                ("sumq", u64 * MAXB), ("nev", u64), ("npsp", u64), ("nbpsw", u64)]


class SieveStats(ctypes.Structure):
    _fields_ = [("primes", u64), ("twins", u64), ("max_gap", u64), ("max_gap_p", u64), ("runs", u64 * 64),
                ("first", u64 * 64)]


STATS_MAXHI = (1 << 64) - (1 << 20)


class Range(ctypes.Structure):
    _fields_ = [("lo", u64), ("hi", u64)]


class Config(ctypes.Structure):
    _fields_ = [("seg_bytes", u32), ("sieve_block", u32), ("test_block", u32), ("arith", u32), ("bucket_min", u32),
                ("chunk", u64), ("test_wgs", u32), ("prove", u32), ("derive", u32), ("fault_step", u32), ("fault_p", u64),
                ("pipeline", u32), ("pad", u32)]


class Timing(ctypes.Structure):
    _fields_ = [("sieve_ms", ctypes.c_double), ("test_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("tested", u64), ("launches", u64), ("plan_ms", ctypes.c_double), ("proved", u64),
                ("psp_ms", ctypes.c_double), ("psp_tests", u64)]


class WFError(RuntimeError):
    def __init__(self, rc, what):
        super().__init__("%s: %s (%d)" % (what, _lib.wf_strerror(rc).decode(), rc))
        self.rc = rc


_lib = None


def load():
    """Load libwf.so (built by mathpub_amd/build.py); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError("libwf.so is not built: run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(LIB)
    P64 = ctypes.POINTER(u64)
    L.wf_strerror.restype = ctypes.c_char_p
    L.wf_strerror.argtypes = [ctypes.c_int]
    L.wf_open.argtypes = [ctypes.c_int, ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
    L.wf_close.argtypes = [ctypes.c_void_p]
    L.wf_close.restype = None
    L.wf_info.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.wf_set_bases.argtypes = [ctypes.c_void_p, ctypes.POINTER(u32), ctypes.c_int]
    L.wf_scan.argtypes = [ctypes.c_void_p, u64, u64, i64, ctypes.POINTER(Result), ctypes.POINTER(Event), u64,
                          ctypes.c_int, P64, u64]
    L.wf_scan_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Range), u32, i64, ctypes.POINTER(Result),
                                ctypes.POINTER(Event), u64]
    L.wf_quotients.argtypes = [ctypes.c_void_p, P64, P64, P64, P64, u64]
    L.wf_quotients_base.argtypes = [ctypes.c_void_p, u32, P64, P64, P64, P64, u64]
    L.wf_quotients_static.argtypes = [ctypes.c_void_p, u32, P64, P64, P64, P64, u64]
    L.wf_primes.argtypes = [ctypes.c_void_p, u64, u64, P64, u64, P64]
    L.wf_bpsw.argtypes = [ctypes.c_void_p, P64, ctypes.POINTER(ctypes.c_uint8), u64]
    L.wf_sieve_stats.argtypes = [ctypes.c_void_p, u64, u64, ctypes.POINTER(SieveStats)]
    L.wf_last_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(Timing)]
    L.wf_autotune.argtypes = [ctypes.c_void_p, u64, u64, ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_double)]
    L.wf_autotune_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Range), u32, ctypes.POINTER(Config),
                                    ctypes.POINTER(ctypes.c_double)]
    L.wf_serve_start.argtypes = [ctypes.c_void_p, u32, u64]
    L.wf_serve_submit.argtypes = [ctypes.c_void_p, u64, u64, i64, P64]
    L.wf_serve_poll.argtypes = [ctypes.c_void_p, u64, ctypes.POINTER(Result), ctypes.POINTER(Event), u64,
                                ctypes.POINTER(ctypes.c_int)]
    L.wf_serve_stop.argtypes = [ctypes.c_void_p]
    for f in ("wf_open", "wf_info", "wf_set_bases", "wf_scan", "wf_scan_batch", "wf_quotients", "wf_quotients_base",
              "wf_quotients_static",
              "wf_primes", "wf_bpsw", "wf_sieve_stats",
              "wf_last_timing", "wf_autotune", "wf_autotune_batch", "wf_serve_start", "wf_serve_submit", "wf_serve_poll", "wf_serve_stop"):
        getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L


def _hi(hi):
    if hi == 1 << 64:
        return 0
    if not 0 <= hi < 1 << 64:
        raise ValueError("hi out of range")
    return hi


def _check(rc, what):
    if rc != 0:
        raise WFError(rc, what)


def result_dict(res, events=None, nb=1):
    """Totals of one range: cks / sumq of the first base, cks_all / sumq_all of every base in
    the engine's order; events as (type, p, q, r, base) sorted by (p, base)."""
    return {"primes": res.primes, "tests": res.tests, "near": res.near, "flt": res.flt,
            "cks": res.cks[0], "sumq": res.sumq[0], "cks_all": list(res.cks[:nb]), "sumq_all": list(res.sumq[:nb]),
            "nev": res.nev, "npsp": res.npsp, "nbpsw": res.nbpsw, "events": events if events is not None else []}


class Engine:
    """One GPU engine (wf_open); not thread-safe, like the C object."""

    def __init__(self, device=0, seg_bytes=0, sieve_block=0, arith=0, chunk=0, bucket_min=0, test_wgs=0,
                 prove=PROVE_AUTO, derive=DERIVE_AUTO, fault_p=0, fault_step=0, pipeline=PIPELINE_GRAPH):
        """fault_p and fault_step are for testing only: the residue of p = fault_p is taken as
        r + 1, so its pair fails the FLT check and is reported as an FLT event; serving step
        number fault_step fails (every request then ends with WF_E_DEVICE)."""
        L = load()
        cfg = Config(seg_bytes, sieve_block, 0, arith, bucket_min, chunk, test_wgs, prove, derive, fault_step, fault_p,
                     pipeline, 0)
        h = ctypes.c_void_p()
        _check(L.wf_open(device, ctypes.byref(cfg), ctypes.byref(h)), "wf_open")
        self._h = h
        self._L = L
        self._bases = [2]
        self._ev = None   # the event array the calls share (events are copied out before a call returns)

    def close(self):
        if self._h:
            self._L.wf_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        buf = ctypes.create_string_buffer(256)
        _check(self._L.wf_info(self._h, buf, 256), "wf_info")
        return buf.value.decode()

    def set_bases(self, bases):
        """Bases of the run, ascending and distinct (wf_set_bases)."""
        arr = (u32 * len(bases))(*bases)
        _check(self._L.wf_set_bases(self._h, arr, len(bases)), "wf_set_bases")
        self._bases = list(bases)

    def _evbuf(self, n):
        n = max(n, 1)
        if self._ev is None or len(self._ev) < n:
            self._ev = (Event * n)()
        return self._ev

    def _events(self, ev, n, with_range=False):
        if with_range:
            return [(e.type, e.p, e.q, (e.r_hi << 64) | e.r_lo, self._bases[e.j], e.range) for e in ev[:n]]
        return [(e.type, e.p, e.q, (e.r_hi << 64) | e.r_lo, self._bases[e.j]) for e in ev[:n]]

    def scan(self, lo, hi, near_T=0, ev_cap=1 << 16, psp=False, psp_cap=1 << 16):
        res = Result()
        ev = self._evbuf(ev_cap)
        pbuf = (u64 * max(psp_cap, 1))()
        _check(self._L.wf_scan(self._h, lo, _hi(hi), near_T, ctypes.byref(res), ev, ev_cap, 1 if psp else 0,
                               pbuf, psp_cap), "wf_scan")
        n = min(res.nev, ev_cap)
        out = result_dict(res, self._events(ev, n), len(self._bases))
        out["psp"] = list(pbuf[:min(res.npsp, psp_cap)]) if psp else None
        return out

    def scan_batch(self, ranges, near_T=0, ev_cap=1 << 16):
        """Per-range totals and the events of all ranges as (type, p, q, r, base, range index),
        sorted by (p, base, range)."""
        n = len(ranges)
        rr = (Range * n)(*[Range(lo, _hi(hi)) for lo, hi in ranges])
        res = (Result * n)()
        ev = self._evbuf(ev_cap)
        _check(self._L.wf_scan_batch(self._h, rr, n, near_T, res, ev, ev_cap), "wf_scan_batch")
        total_ev = sum(r.nev for r in res)
        events = self._events(ev, min(total_ev, ev_cap), with_range=True)
        return [result_dict(r, None, len(self._bases)) for r in res], events

    def quotients(self, ps, base=2):
        """[(q, r)] with r = base^(p-1) mod p^2, q = (r - 1) p^-1 mod 2^64 (wf_quotients_base)."""
        n = len(ps)
        arr = (u64 * n)(*ps)
        q, rh, rl = (u64 * n)(), (u64 * n)(), (u64 * n)()
        _check(self._L.wf_quotients_base(self._h, base, arr, q, rh, rl, n), "wf_quotients_base")
        return [(q[i], (rh[i] << 64) | rl[i]) for i in range(n)]

    def quotients_static(self, ps, base=2):
        """The same through wf_quotients_static (no device allocation; at most STATIC_QUOT_MAX)."""
        n = len(ps)
        arr = (u64 * max(n, 1))(*ps)
        q, rh, rl = (u64 * max(n, 1))(), (u64 * max(n, 1))(), (u64 * max(n, 1))()
        _check(self._L.wf_quotients_static(self._h, base, arr, q, rh, rl, n), "wf_quotients_static")
        return [(q[i], (rh[i] << 64) | rl[i]) for i in range(n)]

    def primes(self, lo, hi):
        cnt = u64()
        _check(self._L.wf_primes(self._h, lo, _hi(hi), None, 0, ctypes.byref(cnt)), "wf_primes")
        buf = (u64 * max(cnt.value, 1))()
        _check(self._L.wf_primes(self._h, lo, _hi(hi), buf, cnt.value, ctypes.byref(cnt)), "wf_primes")
        return list(buf[:cnt.value])

    def bpsw(self, ns):
        """BPSW flags of each n (wf_bpsw): BPSW_SPRP2 | BPSW_LUCAS | BPSW_PRIME bits."""
        n = len(ns)
        arr = (u64 * max(n, 1))(*ns)
        out = (ctypes.c_uint8 * max(n, 1))()
        _check(self._L.wf_bpsw(self._h, arr, out, n), "wf_bpsw")
        return list(out[:n])

    def sieve_stats(self, lo, hi):
        """Primes, twin pairs, largest gap and A158939 run lengths of the primes of [lo, hi)
        (wf_sieve_stats): runs / first map n to the count and the smallest p with L(p) = n."""
        st = SieveStats()
        _check(self._L.wf_sieve_stats(self._h, lo, hi, ctypes.byref(st)), "wf_sieve_stats")
        runs = {n: st.runs[n] for n in range(64) if st.runs[n]}
        first = {n: st.first[n] for n in runs}
        return {"primes": st.primes, "twins": st.twins, "max_gap": st.max_gap, "max_gap_p": st.max_gap_p,
                "runs": runs, "first": first}

    def autotune(self, lo, hi):
        """Pick segment size, sieve workgroup and bucket threshold for [lo, hi) (wf_autotune)."""
        c, ms = Config(), ctypes.c_double()
        _check(self._L.wf_autotune(self._h, lo, _hi(hi), ctypes.byref(c), ctypes.byref(ms)), "wf_autotune")
        return {"seg_bytes": c.seg_bytes, "sieve_block": c.sieve_block, "bucket_min": c.bucket_min,
                "sieve_ms_per_1e9": ms.value}

    def autotune_batch(self, ranges):
        """The same for a batch of ranges (wf_autotune_batch): the first ranges, up to 3e9 numbers."""
        n = len(ranges)
        rr = (Range * n)(*[Range(lo, _hi(hi)) for lo, hi in ranges])
        c, ms = Config(), ctypes.c_double()
        _check(self._L.wf_autotune_batch(self._h, rr, n, ctypes.byref(c), ctypes.byref(ms)), "wf_autotune_batch")
        return {"seg_bytes": c.seg_bytes, "sieve_block": c.sieve_block, "bucket_min": c.bucket_min,
                "sieve_ms_per_1e9": ms.value}

    # serving (continuous batching inside the library; see wf.h)
    def serve_start(self, slots=64, hbm_budget=0):
        _check(self._L.wf_serve_start(self._h, slots, hbm_budget), "wf_serve_start")

    def serve_submit(self, lo, hi, near_T=0):
        """Ticket, or None when the server is full (WF_E_BUSY: retry later)."""
        t = u64()
        rc = self._L.wf_serve_submit(self._h, lo, _hi(hi), near_T, ctypes.byref(t))
        if rc == -6:
            return None
        _check(rc, "wf_serve_submit")
        return t.value

    def serve_poll(self, ticket, ev_cap=1 << 16):
        """Result dict (with its events, as scan returns them) once the request is done, else
        None.  Raises WFError (WF_E_OVERFLOW) for a request whose events overflowed."""
        res, done = Result(), ctypes.c_int()
        ev = self._evbuf(ev_cap)
        _check(self._L.wf_serve_poll(self._h, ticket, ctypes.byref(res), ev, ev_cap, ctypes.byref(done)),
               "wf_serve_poll")
        if not done.value:
            return None
        return result_dict(res, self._events(ev, min(res.nev, ev_cap)), len(self._bases))

    def serve_stop(self):
        _check(self._L.wf_serve_stop(self._h), "wf_serve_stop")

    def last_timing(self):
        t = Timing()
        _check(self._L.wf_last_timing(self._h, ctypes.byref(t)), "wf_last_timing")
        return {"sieve_ms": t.sieve_ms, "test_ms": t.test_ms, "total_ms": t.total_ms, "tested": t.tested,
                "launches": t.launches, "plan_ms": t.plan_ms, "proved": t.proved, "psp_ms": t.psp_ms,
                "psp_tests": t.psp_tests}
