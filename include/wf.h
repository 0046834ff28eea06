/*
 * wf.h -- C ABI of the MI355X Wieferich search engine (libwf.so).
 *
 * The engine runs the search loop of A001220/wieferich.c (jeffsponaugle/Mathpub) on one
 * AMD MI355X: a wheel-30 segmented sieve over [lo, hi) on the GPU, then for every odd
 * prime p and every base b of the run the Fermat quotient q = (b^(p-1) mod p^2 - 1) / p,
 * its checksum term mix64(q ^ p * 0x9E3779B97F4A7C15), the FLT check q < p and the
 * FOUND / NEAR events.
 * Results are bit-identical to the reference CPU tool.
 *
 * The reference has no plugin API; the entry points below replace its two internal seams
 * (the per-chunk scan and the prime source):
 *   wf_scan        scan_chunk()          A001220/wieferich.c:1129-1177 (with the prime source,
 *                  primesieve_jump_to/next_prime, :1138-1141) and the GPU tool's
 *                  sieve_chunk/enqueue_chunk/fold_chunk, A001220/cuda/wieferich_cuda.cu:1090-1271
 *   wf_scan_batch  (no counterpart: many uneven ranges in one launch)
 *   wf_quotients   fermat_quotient()     A001220/wieferich.c:939-949, GPU k_quot
 *   (_base)        A001220/cuda/wieferich_cuda.cu:976-983 (check / selftest)
 *   wf_primes      primesieve_jump_to / primesieve_next_prime, A001220/wieferich.c:1138-1141
 *                  (the prime set of [lo, hi), for tests and the sieve benchmark)
 *   wf_bpsw        d_sprp2_64 + d_slprp64 of A350826/cuda/a350826_cuda.cu:768-898 (its
 *                  test_kernel / `num` command, :2458-2483)
 *   wf_sieve_stats the GPU sieve for the repository's other prime consumers: twin counts
 *                  (A007508/README.md:3-5, 44-64) and runs of increasing gaps
 *                  (A158939/a158939.c:19-24, A158939/cuda/a158939_cuda.cu:53-78)
 *   wf_serve_*     (no counterpart: continuous batching of range queries)
 *
 * Conventions (as wieferich.c): return 0 on success or a negative WF_E_* code
 * (wf_strerror names it; the host program prints it and exits with status 1, as die()
 * does, wieferich.c:271-280).  A range is [lo, hi) with hi == 0 meaning 2^64.  The
 * prime 2 is counted but not tested (wieferich.c:1142-1146).  An engine belongs to one
 * host thread; calls on one engine must not overlap.  The caller owns every output
 * buffer; the engine owns its device memory.  Events beyond ev_cap are counted in nev
 * and never silently dropped (the caller sees nev > ev_cap).
 */
#ifndef WF_H
#define WF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_MAXB 64

/* Opaque engine: one per GPU, owned by one host thread. */
typedef struct wf_engine wf_engine;

enum {
    WF_OK = 0,
    WF_E_ARG = -1,          /* bad argument (range, sizes, NULL pointer) */
    WF_E_DEVICE = -2,       /* no GPU, or a HIP runtime call failed */
    WF_E_NOMEM = -3,        /* device or host allocation failed */
    WF_E_UNSUPPORTED = -4,  /* a feature the device cannot run (segments above its LDS) */
    WF_E_OVERFLOW = -5,     /* an internal buffer bound was exceeded (bug guard) */
    WF_E_BUSY = -6,         /* serving: no free slot / HBM budget; or a call that needs the
                               engine (scans, quotients, BPSW, statistics, autotuning,
                               timings, bases) while it serves, until wf_serve_stop */
    WF_E_LIMIT = -7,        /* wf_quotients_static: n above WF_STATIC_QUOT_MAX */
};

/* Event types, the reference's result lines (wieferich.c:1009, :1060, :1071). */
enum { WF_EV_FOUND = 0, WF_EV_NEAR = 1, WF_EV_FLT = 2 };

typedef struct {
    uint32_t type, j;       /* type, base index */
    uint64_t p, q;          /* prime and Fermat quotient (q >= p for FLT) */
    uint64_t r_hi, r_lo;    /* b^(p-1) mod p^2 */
    uint32_t range;         /* index of the range in wf_scan_batch's list (0 for wf_scan and
                               for the events of a served request) */
    uint32_t pad;
} wf_event;

/* Totals of one scanned range: the reference's slot_t (wieferich.c:956-960) plus the
 * sum of the quotients and the event / pseudoprime counts. */
typedef struct {
    uint64_t primes, tests, near, flt;
    uint64_t cks[WF_MAXB];
    uint64_t sumq[WF_MAXB];
    uint64_t nev;           /* events of this range (all of them, even beyond ev_cap) */
    uint64_t npsp;          /* odd composites that are base-2 strong probable primes */
    uint64_t nbpsw;         /* of those, the ones that also pass the strong Lucas test (BPSW
                               pseudoprimes: none exist below 2^64) */
} wf_result;

typedef struct { uint64_t lo, hi; } wf_range;   /* [lo, hi), hi == 0 means 2^64 */

/* Tuning knobs; a zero field means "engine default". */
typedef struct {
    uint32_t seg_bytes;     /* LDS sieve segment, 4096 .. 131072 (a power of two) */
    uint32_t sieve_block;   /* threads per sieve workgroup (64 .. 1024, a multiple of 64) */
    uint32_t test_block;    /* threads per test workgroup: 256, the only size the test
                               kernels are built for (others give WF_E_ARG) */
    uint32_t arith;         /* 0 = fastest exact path, 1 = force 32-bit Montgomery */
    uint32_t bucket_min;    /* sieving primes >= this go through bucket records (>= 512) */
    uint64_t chunk;         /* numbers per device pass of wf_scan (memory bound) */
    uint32_t test_wgs;      /* wf_scan_batch: test workgroups (4 wavefronts) per CU, 1 .. 16;
                               fewer leave room for the next pass's sieve beside the tests */
    uint32_t prove;         /* partial sieve + proof for short pieces: WF_PROVE_AUTO (by a cost
                               model), WF_PROVE_OFF, WF_PROVE_ON (every piece above bucket_min^2) */
    uint32_t derive;        /* composite bases from their prime factors' quotients (p < 9 * 2^47):
                               WF_DERIVE_AUTO (whenever that needs fewer ladders than bases),
                               WF_DERIVE_OFF (one ladder per base) */
    uint32_t fault_step;    /* testing only, 0 = off: serving step number fault_step (counted
                               from 1) fails with WF_E_DEVICE instead of being enqueued */
    uint64_t fault_p;       /* testing only, 0 = off: the residue of this p is taken as r + 1,
                               which fails the FLT check (exercises the FLT-ERROR path: the
                               reference's finish_lanes, wieferich.c:1058-1063) */
    uint32_t pipeline;      /* wf_scan_batch: WF_PIPELINE_GRAPH (all passes in one HIP graph, the
                               sieve of pass k + 1 beside the tests of pass k) or WF_PIPELINE_OFF
                               (the passes one after another on one stream, each kernel timed:
                               the plain engine, as wf_scan runs a range) */
    uint32_t pad;
} wf_config;

enum { WF_PIPELINE_GRAPH = 0, WF_PIPELINE_OFF = 1 };

/* Prove mode: a piece is sieved by the primes below bucket_min only, and its survivors are
 * decided by BPSW (base-2 strong test + strong Lucas test, exact below 2^64) instead of
 * placing every sieving prime up to sqrt(hi) in it.  Results are identical either way. */
enum { WF_PROVE_AUTO = 0, WF_PROVE_OFF = 1, WF_PROVE_ON = 2 };

/* Composite bases: the Fermat quotient is additive, q_p(ab) = q_p(a) + q_p(b) (mod p) for
 * p not dividing ab, so the quotients of every base follow from ladders for the distinct
 * primes dividing the bases (the 17 "oeis" bases need 9).  Results are identical either way. */
enum { WF_DERIVE_AUTO = 0, WF_DERIVE_OFF = 1 };

/* Engine life cycle (the reference's gpu_init, wieferich_cuda.cu:1570-1577). */
int  wf_open(int device, const wf_config *cfg, wf_engine **out);
void wf_close(wf_engine *e);
const char *wf_strerror(int rc);
/* Device name, CU count and the arithmetic path, for status lines. */
int  wf_info(wf_engine *e, char *buf, int len);

/* Bases of the run, ascending and distinct, 2 <= b < 2^32, at most WF_MAXB (the -b list of
 * parse_bases, wieferich.c:816-849).  Results carry one checksum and one sum of q per base
 * (index j = position in this list); pairs with p | b are skipped, as process_batch does
 * (wieferich.c:1097).  Base 2 runs its own ladders; any other base uses windows of w bits
 * with multipliers b^d < 2^32 (wieferich.c:99-107).  wf_config.arith = 1 computes every base
 * with 32-bit limb Montgomery ladders: doublings for base 2, the binary ladder (a product by
 * b R mod p^2 per set bit) for the others.  The default is the single base 2. */
int  wf_set_bases(wf_engine *e, const uint32_t *bases, int nb);

/* Scan one range: totals into *out, events sorted by (p, j) into ev[0..ev_cap).
 * near_T > 0 reports 0 < |A| <= near_T (wieferich.c:1065-1073).  want_psp != 0 also
 * lists the odd composites of the range that are base-2 strong probable primes, sorted,
 * into psp[0..psp_cap) (out->npsp counts all of them). */
int  wf_scan(wf_engine *e, uint64_t lo, uint64_t hi, int64_t near_T, wf_result *out,
             wf_event *ev, uint64_t ev_cap, int want_psp, uint64_t *psp, uint64_t psp_cap);

/* Scan n ranges in one device pass: per_range[i] receives range i's totals; events of
 * all ranges go to ev, sorted by (p, j, range), each with the index of its range (ranges may
 * overlap or repeat: an event of a number in two ranges appears once per range). */
int  wf_scan_batch(wf_engine *e, const wf_range *r, uint32_t n, int64_t near_T,
                   wf_result *per_range, wf_event *ev, uint64_t ev_cap);

/* Fermat quotients of arbitrary odd p (prime or not) for base 2: r = 2^(p-1) mod p^2,
 * q = (r - 1) p^-1 mod 2^64, as fermat_quotient() (wieferich.c:939-949). */
int  wf_quotients(wf_engine *e, const uint64_t *p, uint64_t *q, uint64_t *r_hi,
                  uint64_t *r_lo, uint64_t n);
/* The same for any base 2 <= b < 2^32 (the check command, wieferich.c:1596-1650; r = 0
 * when p | b). */
int  wf_quotients_base(wf_engine *e, uint32_t b, const uint64_t *p, uint64_t *q,
                       uint64_t *r_hi, uint64_t *r_lo, uint64_t n);

/* wf_quotients_base for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  It calls no allocating routine: its staging area is a static device array of
 * WF_STATIC_QUOT_MAX values per GPU (8 MiB, held until the process ends), so n above that
 * returns WF_E_LIMIT.  Every call on one GPU shares the array: calls that overlap on one GPU
 * race. */
#define WF_STATIC_QUOT_MAX (1u << 18)
int  wf_quotients_static(wf_engine *e, uint32_t b, const uint64_t *p, uint64_t *q,
                         uint64_t *r_hi, uint64_t *r_lo, uint64_t n);

/* The primes of [lo, hi) in ascending order into out[0..cap); *count = how many exist. */
int  wf_primes(wf_engine *e, uint64_t lo, uint64_t hi, uint64_t *out, uint64_t cap,
               uint64_t *count);

/* Statistics of the primes of [lo, hi) for the repository's other sieve consumers.
 * L(p) is A158939's run length: with q_0 = p < q_1 < q_2 < ... the primes from p on and
 * g_i = q_i - q_(i-1), L(p) = the largest n with g_1 < g_2 < ... < g_n (always >= 1). */
typedef struct {
    uint64_t primes;            /* primes p in [lo, hi) */
    uint64_t twins;             /* p in [lo, hi) with p + 2 prime (A007508: pairs below hi) */
    uint64_t max_gap;           /* largest gap q_1 - p over p in [lo, hi) */
    uint64_t max_gap_p;         /* the smallest p with that gap */
    uint64_t runs[64];          /* runs[n] = primes p in [lo, hi) with L(p) = n */
    uint64_t first[64];         /* the smallest of them (0 if none) */
} wf_stats;

/* The gaps and runs after a prime near hi use the primes beyond it: hi <= WF_STATS_MAXHI. */
#define WF_STATS_MAXHI (0xFFFFFFFFFFFFFFFFull - (1ull << 20) + 1)
int  wf_sieve_stats(wf_engine *e, uint64_t lo, uint64_t hi, wf_stats *out);

/* BPSW verdicts of n[0..count): flags[i] bit 0 = base-2 strong probable prime, bit 1 =
 * strong Lucas probable prime with Selfridge's parameters (both for odd n >= 3), bit 2 =
 * prime (exact below 2^64).  Any n is accepted (even n and n < 3 give 0, or 4 for n = 2). */
enum { WF_BPSW_SPRP2 = 1, WF_BPSW_LUCAS = 2, WF_BPSW_PRIME = 4 };
int  wf_bpsw(wf_engine *e, const uint64_t *n, uint8_t *flags, uint64_t count);

/* Serving: continuous batching over the device-resident sieving-prime table.  A server
 * thread owns the engine between wf_serve_start and wf_serve_stop.  It runs steps: each takes
 * the next piece of every admitted request and sieves and tests all of them in the same
 * kernel launches.  It plans and enqueues step k + 1 while step k runs on the GPU, so requests
 * that arrive meanwhile join the next step.
 * submit admits a range with its near-hit bound when a slot and the HBM budget allow (else
 * WF_E_BUSY: retry later).  poll returns its totals once done (*done = 1), with its FOUND /
 * NEAR / FLT events sorted by (p, j) into ev[0..ev_cap) (out->nev counts all of them).  The
 * results equal wf_scan's of the same range and near_T.  A request whose events did not all
 * fit the step's device event buffer is returned with WF_E_OVERFLOW (and *done = 1); it is
 * never truncated silently.  A failed step ends every request with its error code, and the
 * server takes no more until wf_serve_stop. */
int  wf_serve_start(wf_engine *e, uint32_t slots, uint64_t hbm_budget);
int  wf_serve_submit(wf_engine *e, uint64_t lo, uint64_t hi, int64_t near_T, uint64_t *ticket);
int  wf_serve_poll(wf_engine *e, uint64_t ticket, wf_result *out, wf_event *ev, uint64_t ev_cap, int *done);
int  wf_serve_stop(wf_engine *e);

/* Autotuner: times the sieve stage on a sample of [lo, hi) over segment sizes 16-128 KiB,
 * sieve workgroups of 4-16 wavefronts and bucket thresholds 2^15-2^20, switches the engine
 * to the fastest and returns it (*best_ms = its sieve milliseconds per 1e9 numbers).
 * No counterpart in the reference (its chunk and thread counts are set by hand). */
int  wf_autotune(wf_engine *e, uint64_t lo, uint64_t hi, wf_config *best, double *best_ms);
/* The same for a batch of ranges (wf_scan_batch's workload): the sample is the ranges in
 * order, up to 3e9 numbers, sieved as one packed pass. */
int  wf_autotune_batch(wf_engine *e, const wf_range *r, uint32_t n, wf_config *best, double *best_ms);

/* Timings of the last wf_scan / wf_scan_batch, or of the serving session ended by the last
 * wf_serve_stop: device milliseconds of the sieve and the test kernels (HIP events on the
 * streams they run on; a pipelined batch times its whole graph only, total_ms), primes
 * tested, test launches, the host milliseconds spent planning the passes, the primes that
 * came through prove mode, and the strong-pseudoprime kernel's time and test count. */
typedef struct { double sieve_ms, test_ms, total_ms; uint64_t tested; uint64_t launches; double plan_ms;
                 uint64_t proved;   /* primes found by the prove-mode path */
                 double psp_ms;     /* the strong-pseudoprime kernel (wf_scan with want_psp) */
                 uint64_t psp_tests;   /* base-2 strong tests it ran (numbers its filters kept) */
               } wf_timing;
int  wf_last_timing(wf_engine *e, wf_timing *t);

#ifdef __cplusplus
}
#endif
#endif
