/*
 * wf_main.c -- the wf command: a drop-in for A001220/wieferich.c's command line, running
 * the scan on one MI355X through libwf.so (include/wf.h).
 *
 *   wf scan [START] END [-b BASES] [-c CHUNK] [-n T] [-S FILE] [-i SECS] [-L FILE] [-g DEV]
 *           [-q] [-Q] [--prove auto|off|on] [--fault-p P]
 *   wf check P [-b BASES] [-g DEV]
 *   wf bench [N] [-d SPAN] [-b BASES] [-g DEV]
 *   wf selftest [LIMIT] [-g DEV]
 *   wf num X [X ...] [-g DEV]
 *   wf dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID] [-g DEV]
 *   wf stats [LO] HI [-g DEV]
 *
 * What a user of the reference sees unchanged (its interface, wieferich.c's Usage): the
 * commands and options, the number and base-list grammar, END rounded up to a chunk
 * boundary (:1389-1394), the chunk log, the checkpoint file (resumable by either tool), the
 * FOUND / NEAR / FLT-ERROR / CONFIRMED lines (:1009, :1060, :1071, :1350-1352), SUMMARY
 * (:1505-1523) and the exit statuses: 0, 1 for an error (die, :271-280), 2 for a usage error
 * (:1535-1544), 3 when any FLT check failed (:1591).  The text formats live in wf_text.c.
 * Differences: the default CHUNK is 10^10, as in the reference's GPU tool
 * (wieferich_cuda.cu:1591); END may be 2^64; -Q adds the sum of the quotients; -t and -T
 * are accepted and ignored (there are no host sieve threads); --fault-p (testing only)
 * corrupts one residue on the GPU to exercise the FLT-ERROR path; num, dcheck and stats are
 * the other tools' commands on the same engine.
 *
 * The chunks of a scan are run one after another, each a wf_scan call, so the frontier
 * (the chunks complete from START on) is simply the count of chunks done.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <inttypes.h>
#include <limits.h>
#include <math.h>
#include <signal.h>
#include <stdarg.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "wf.h"
#include "wf_text.h"

typedef uint64_t u64;
typedef uint32_t u32;
typedef wt_u128 u128;

#define EVENT_CAP (1u << 16)     /* events one chunk may report (FOUND / NEAR / FLT) */

static bool g_tty;                          /* stderr is a terminal: the status line is redrawn */
static volatile sig_atomic_t g_interrupted;
static wf_engine *g_gpu;
static int g_device;
static wf_config g_conf;                    /* zero fields: the engine's defaults */
static u32 g_base[WF_MAXB];
static int g_nbase;

/* ------------------------------------------------------------------ errors, arguments */

static void clear_status(void)
{
    if (g_tty) fputs("\r\033[K", stderr);
}

static void fail(const char *fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));
static void fail(const char *fmt, ...)
{
    va_list ap;
    clear_status();
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    exit(1);
}

static void usage_exit(void) __attribute__((noreturn));
static void usage_exit(void)
{
    fputs("usage: wf scan [START] END [-b BASES] [-c CHUNK] [-n T] [-S FILE] [-i SECS] [-L FILE] [-g DEV] [-q] [-Q]\n"
          "       wf check P [-b BASES] [-g DEV]\n"
          "       wf bench [N] [-d SPAN] [-b BASES] [-g DEV]\n"
          "       wf selftest [LIMIT] [-g DEV]\n"
          "       wf num X [X ...] [-g DEV]          BPSW tests of single numbers on the GPU\n"
          "       wf dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID] [-g DEV]   recompute every STEP-th chunk\n"
          "       wf stats [LO] HI [-g DEV]          primes, twin pairs, largest gap, runs of increasing gaps\n"
          "scan --prove auto|off|on: sieve short chunks with small primes only and prove the rest by BPSW\n"
          "scan --fault-p P (testing only): corrupt the residue of P, which must then be reported as FLT-ERROR\n"
          "BASES: list/ranges like 3,5,7 or 2-30; 'oeis' (default) = 3,5,6,7,10,12,13,14,15,17,18,19,20,22,23,26,30;\n"
          "       'known' = the bases 2..30 that are not perfect powers (default of check)\n"
          "END may be 2^64\n", stderr);
    exit(2);
}

static u128 num128(const char *s)
{
    u128 v = 0;
    switch (wt_number(s, &v)) {
        case WT_OK: return v;
        case WT_OVERFLOW: fail("overflow in '%s'", s);
        case WT_NEGATIVE: fail("negative result in '%s'", s);
        default: fail("bad number '%s'", s);
    }
}

static u64 num64(const char *s)
{
    const u128 v = num128(s);
    if (v >= WT_TWO64) fail("overflow in '%s'", s);
    return (u64)v;
}

static int num_in(const char *s, int lo, int hi, const char *what)
{
    const u64 v = num64(s);
    if (v < (u64)lo || v > (u64)hi) fail("%s must be between %d and %d", what, lo, hi);
    return (int)v;
}

/* An option that takes a value: argv[*i] names it and argv[*i + 1] holds it. */
static bool opt(int argc, char **argv, int *i, const char *name, const char **val)
{
    if (strcmp(argv[*i], name) != 0 || *i + 1 >= argc) return false;
    *val = argv[++*i];
    return true;
}

/* Options every command takes: -g DEV, and -t / -T (threads of the CPU tool; ignored). */
static bool common_opt(int argc, char **argv, int *i)
{
    const char *v;
    if (opt(argc, argv, i, "-g", &v)) g_device = num_in(v, 0, 64, "device");
    else if (opt(argc, argv, i, "-t", &v) || opt(argc, argv, i, "-T", &v)) (void)num_in(v, 1, 1024, "threads");
    else return false;
    return true;
}

static void choose_bases(const char *spec)
{
    char err[256];
    if (!wt_bases(spec, g_base, &g_nbase, err, sizeof err)) fail("%s", err);
    if (g_gpu) {
        const int rc = wf_set_bases(g_gpu, g_base, g_nbase);
        if (rc) fail("wf_set_bases: %s", wf_strerror(rc));
    }
}

static void open_gpu(void)
{
    if (g_gpu) return;
    int rc = wf_open(g_device, &g_conf, &g_gpu);
    if (rc) fail("cannot open GPU %d: %s", g_device, wf_strerror(rc));
    if ((rc = wf_set_bases(g_gpu, g_base, g_nbase))) {
        char b[1024];
        wt_join_bases(b, sizeof b, g_base, g_nbase);
        fail("bases %s: %s", b, wf_strerror(rc));
    }
}

static void close_gpu(void)
{
    if (g_gpu) wf_close(g_gpu);
    g_gpu = NULL;
}

static double seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void on_signal(int sig)
{
    (void)sig;
    g_interrupted = 1;
}

/* ------------------------------------------------------------------ a scan */

typedef struct {
    u64 start, chunk, nchunks, frontier;
    u128 end;                       /* rounded up to a chunk boundary; may be 2^64 */
    u64 primes, tests, near, flt;
    u64 cks[WF_MAXB], sumq[WF_MAXB];
    wt_hit *hits;
    size_t nhits, hcap;
} scan_t;

typedef struct {
    long long near_T;               /* -n */
    const char *state;              /* -S */
    FILE *log;                      /* -L */
    int interval;                   /* -i */
    bool quiet;                     /* -q */
    bool print;                     /* FOUND / NEAR / CONFIRMED lines */
} scan_opts;

typedef struct { double seconds, gpu_ms; u64 primes, tests; u128 numbers; } session_t;

static u64 chunk_start(const scan_t *S, u64 i) { return S->start + i * S->chunk; }

static u128 chunk_end(const scan_t *S, u64 i)
{
    const u128 e = (u128)S->start + (u128)(i + 1) * S->chunk;
    return e < S->end ? e : S->end;
}

/* Position of the frontier: every prime below it has been tested. */
static u128 frontier_pos(const scan_t *S)
{
    const u128 p = (u128)S->start + (u128)S->frontier * S->chunk;
    return p < S->end ? p : S->end;
}

static void print_anum_suffix(FILE *f, u32 b)
{
    const char *a = wt_anum(b);
    if (a) fprintf(f, " (%s)", a);
}

static bool known_hit(const scan_t *S, u32 b, u64 p)
{
    for (size_t i = 0; i < S->nhits; i++)
        if (S->hits[i].b == b && S->hits[i].p == p) return true;
    return false;
}

static void keep_hit(scan_t *S, u32 b, u64 p, u64 chunk)
{
    if (S->nhits == S->hcap) {
        S->hcap = S->hcap ? 2 * S->hcap : 16;
        if (!(S->hits = realloc(S->hits, S->hcap * sizeof *S->hits))) fail("out of memory");
    }
    S->hits[S->nhits++] = (wt_hit){b, p, chunk};
}

static void confirm(const scan_t *S, const wt_hit *h)
{
    char pos[24];
    wt_u128_str(pos, sizeof pos, frontier_pos(S));
    clear_status();
    printf("CONFIRMED base %u p %" PRIu64, h->b, h->p);
    print_anum_suffix(stdout, h->b);
    printf(" (all primes in [%" PRIu64 ", %s) tested)\n", S->start, pos);
}

/* The result lines of one chunk's events, which come sorted by (p, base). */
static void report_events(scan_t *S, u64 idx, const wf_event *ev, u64 n, const scan_opts *o)
{
    for (u64 k = 0; k < n; k++) {
        const wf_event *e = &ev[k];
        const u32 b = g_base[e->j];
        switch (e->type) {
            case WF_EV_FOUND:
                if (known_hit(S, b, e->p)) break;
                keep_hit(S, b, e->p, idx);
                if (o->print) {
                    clear_status();
                    printf("FOUND base %u p %" PRIu64, b, e->p);
                    print_anum_suffix(stdout, b);
                    putchar('\n');
                }
                break;
            case WF_EV_NEAR:
                if (o->print) {     /* the symmetric residue A of q: q or q - p */
                    const long long A = e->q <= e->p / 2 ? (long long)e->q : -(long long)(e->p - e->q);
                    clear_status();
                    printf("NEAR base %u p %" PRIu64 " A %lld\n", b, e->p, A);
                }
                break;
            default:                /* an FLT error is always reported */
                clear_status();
                printf("FLT-ERROR base %u p %" PRIu64 " r = %016" PRIx64 "%016" PRIx64 "\n", b, e->p, e->r_hi, e->r_lo);
                break;
        }
    }
    fflush(stdout);
}

/* Fold a finished chunk into the totals and the chunk log; advance the frontier. */
static void fold(scan_t *S, u64 idx, const wf_result *r, const scan_opts *o)
{
    S->primes += r->primes;
    S->tests += r->tests;
    S->near += r->near;
    S->flt += r->flt;
    for (int j = 0; j < g_nbase; j++) {
        S->cks[j] += r->cks[j];
        S->sumq[j] += r->sumq[j];
    }
    if (o->log) {
        char line[64 + WF_MAXB * 32];
        wt_chunk_line(line, sizeof line, chunk_start(S, idx), chunk_end(S, idx), r->primes, g_base, r->cks, g_nbase);
        fprintf(o->log, "%s\n", line);
        fflush(o->log);
    }
    S->frontier = idx + 1;
}

static void save_checkpoint(const scan_t *S, const char *path)
{
    char *tmp = NULL;
    if (asprintf(&tmp, "%s.tmp", path) < 0) fail("out of memory");
    FILE *f = fopen(tmp, "w");
    if (!f) {
        fprintf(stderr, "\nwarning: cannot write %s: %s\n", tmp, strerror(errno));
        free(tmp);
        return;
    }
    wt_ckpt c;
    memset(&c, 0, sizeof c);
    wt_join_bases(c.bases, sizeof c.bases, g_base, g_nbase);
    c.start = S->start;
    c.chunk = S->chunk;
    c.end = S->end;
    c.frontier = S->frontier;
    c.primes = S->primes;
    c.tests = S->tests;
    c.near = S->near;
    c.flt = S->flt;
    c.ncks = g_nbase;
    for (int j = 0; j < g_nbase; j++) {
        c.cks_base[j] = g_base[j];
        c.cks[j] = S->cks[j];
    }
    c.hits = S->hits;
    c.nhits = S->nhits;
    bool ok = wt_ckpt_write(f, &c);
    if (fclose(f) != 0) ok = false;
    if (!ok || rename(tmp, path) != 0) {
        fprintf(stderr, "\nwarning: cannot update %s: %s (previous checkpoint kept)\n", path, strerror(errno));
        remove(tmp);
    }
    free(tmp);
}

/* Resume from `path` if it exists; the run's bases, START, CHUNK and chunk count are set. */
static bool restore_checkpoint(scan_t *S, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        if (errno == ENOENT) return false;
        fail("cannot read %s: %s", path, strerror(errno));
    }
    wt_ckpt c;
    char err[1024 + 256];
    const int rc = wt_ckpt_read(f, path, &c, err, sizeof err);
    fclose(f);
    if (rc) fail("%s", err);
    char cur[WT_BASES_MAX + 1];
    wt_join_bases(cur, sizeof cur, g_base, g_nbase);
    if (strcmp(c.bases, cur) != 0) fail("%s was written for bases %s, not %s", path, c.bases, cur);
    bool have[WF_MAXB] = {false};
    for (int k = 0; k < c.ncks; k++) {
        int j = 0;
        while (j < g_nbase && g_base[j] != c.cks_base[k]) j++;
        if (j == g_nbase) fail("%s: base %u is not in the current base list", path, c.cks_base[k]);
        S->cks[j] = c.cks[k];
        have[j] = true;
    }
    for (int j = 0; j < g_nbase; j++)
        if (!have[j]) fail("%s has no checksum for base %u", path, g_base[j]);
    if (c.start != S->start || c.chunk != S->chunk)
        fail("%s was written for start=%" PRIu64 " chunk=%" PRIu64 ", not start=%" PRIu64 " chunk=%" PRIu64, path,
             c.start, c.chunk, S->start, S->chunk);
    if (c.frontier > S->nchunks)
        fail("%s is past the new END (frontier %" PRIu64 " > %" PRIu64 " chunks)", path, c.frontier, S->nchunks);
    /* END itself may differ: a resumed scan can be extended or shortened */
    S->frontier = c.frontier;
    S->primes = c.primes;
    S->tests = c.tests;
    S->near = c.near;
    S->flt = c.flt;
    for (size_t i = 0; i < c.nhits; i++)
        if (!known_hit(S, c.hits[i].b, c.hits[i].p)) keep_hit(S, c.hits[i].b, c.hits[i].p, c.hits[i].chunk);
    wt_ckpt_free(&c);
    return true;
}

static void status_line(const scan_t *S, double t0, double t, u64 frontier0, u64 tests0, const scan_opts *o)
{
    char el[32], eta_s[32];
    const double span = t - t0, done = (double)(S->frontier - frontier0) * (double)S->chunk;
    const double rate = span > 0 ? done / span : 0, trate = span > 0 ? (double)(S->tests - tests0) / span : 0;
    const double left = (double)(S->nchunks - S->frontier) * (double)S->chunk;
    size_t confirmed = 0;
    for (size_t i = 0; i < S->nhits; i++) confirmed += S->hits[i].chunk < S->frontier;
    wt_duration(el, sizeof el, span);
    wt_duration(eta_s, sizeof eta_s, rate > 0 ? left / rate : -1);
    fprintf(stderr, "%s%s  pos %.6g (%.2f%%)  %.3g/s  %.3g tests/s  %.4g primes  ETA %s  found %zu (%zu confirmed)",
            g_tty ? "\r\033[K" : "progress: ", el, (double)chunk_start(S, S->frontier),
            100.0 * (double)S->frontier / (double)S->nchunks, rate, trate, (double)S->primes, eta_s, S->nhits, confirmed);
    if (o->near_T > 0) fprintf(stderr, "  near %" PRIu64, S->near);
    if (S->flt) fprintf(stderr, "  FLT-ERRORS %" PRIu64, S->flt);
    if (!g_tty) fputc('\n', stderr);
    fflush(stderr);
}

/* Scan [start, end) chunk by chunk (the reference's run_scan, wieferich.c:1382-1495). */
static session_t run_scan(scan_t *S, u64 start, u128 end, u64 chunk, const scan_opts *o)
{
    if (end <= start) fail("END must be larger than START");
    if (chunk == 0) fail("CHUNK must be positive");
    free(S->hits);
    memset(S, 0, sizeof *S);
    S->start = start;
    S->chunk = chunk;
    const u128 n = (end - start + chunk - 1) / chunk;
    S->nchunks = (u64)n;
    S->end = (u128)start + n * chunk;
    if (S->end > WT_TWO64) fail("END must be <= 2^64 (after rounding up to the chunk size)");
    char es[24];
    wt_u128_str(es, sizeof es, S->end);
    if (S->end != end && !o->quiet) fprintf(stderr, "note: END rounded up to the chunk boundary %s\n", es);
    if (o->state && restore_checkpoint(S, o->state) && !o->quiet)
        fprintf(stderr, "resuming from %s: %" PRIu64 " of %" PRIu64 " chunks done (position %" PRIu64 ", %" PRIu64
                        " primes tested)\n", o->state, S->frontier, S->nchunks, chunk_start(S, S->frontier), S->primes);
    if (o->print)                   /* solutions restored from the checkpoint */
        for (size_t i = 0; i < S->nhits; i++) confirm(S, &S->hits[i]);
    const u64 frontier0 = S->frontier, primes0 = S->primes, tests0 = S->tests;
    open_gpu();
    if (!o->quiet) {
        char info[256], bl[1024];
        wf_info(g_gpu, info, sizeof info);
        wt_join_bases(bl, sizeof bl, g_base, g_nbase);
        fprintf(stderr, "GPU %d: %s\nscanning [%" PRIu64 ", %s) in %" PRIu64 " chunks of %" PRIu64 ", bases %s\n",
                g_device, info, S->start, es, S->nchunks, S->chunk, bl);
    }
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_handler = on_signal;
    sigaction(SIGINT, &sa, NULL);
    sigaction(SIGTERM, &sa, NULL);
    wf_event *ev = malloc(EVENT_CAP * sizeof *ev);
    if (!ev) fail("out of memory");
    const double t0 = seconds();
    double shown = t0, saved = t0, gpu_ms = 0;
    while (S->frontier < S->nchunks && !g_interrupted) {
        const u64 idx = S->frontier, lo = chunk_start(S, idx);
        const u128 hi = chunk_end(S, idx);
        wf_result r;
        const int rc = wf_scan(g_gpu, lo, hi >= WT_TWO64 ? 0 : (u64)hi, o->near_T, &r, ev, EVENT_CAP, 0, NULL, 0);
        if (rc == WF_E_OVERFLOW && r.nev > EVENT_CAP)
            fail("chunk [%" PRIu64 ", ...): %" PRIu64 " events exceed the event buffer (%u); use a smaller -n or CHUNK", lo,
                 r.nev, EVENT_CAP);
        if (rc) fail("chunk starting at %" PRIu64 ": %s", lo, wf_strerror(rc));
        wf_timing tm;
        if (wf_last_timing(g_gpu, &tm) == WF_OK) gpu_ms += tm.total_ms;
        const size_t before = S->nhits;
        report_events(S, idx, ev, r.nev < EVENT_CAP ? r.nev : EVENT_CAP, o);
        fold(S, idx, &r, o);
        if (o->print)               /* a chunk's solutions are confirmed once it is folded */
            for (size_t i = before; i < S->nhits; i++) confirm(S, &S->hits[i]);
        fflush(stdout);
        const double t = seconds();
        if (!o->quiet && t - shown >= (g_tty ? 1.0 : (double)o->interval)) {
            shown = t;
            status_line(S, t0, t, frontier0, tests0, o);
        }
        if (o->state && t - saved >= o->interval) {
            saved = t;
            save_checkpoint(S, o->state);
        }
    }
    const double t1 = seconds();
    free(ev);
    if (!o->quiet) clear_status();
    if (o->state) save_checkpoint(S, o->state);
    if (g_interrupted && S->frontier < S->nchunks && !o->quiet)
        fprintf(stderr, "stopped at position %" PRIu64 " (%" PRIu64 " of %" PRIu64 " chunks complete)%s\n",
                chunk_start(S, S->frontier), S->frontier, S->nchunks, o->state ? ", checkpoint saved" : "");
    session_t st;
    st.seconds = t1 - t0;
    st.gpu_ms = gpu_ms;
    st.primes = S->primes - primes0;
    st.tests = S->tests - tests0;
    st.numbers = (u128)(S->frontier - frontier0) * S->chunk;
    return st;
}

static void summary(const scan_t *S, long long near_T, bool with_sumq)
{
    wt_summary(stdout, S->start, frontier_pos(S), S->primes, S->tests, S->flt, near_T, S->near, g_base, g_nbase, S->cks,
               with_sumq ? S->sumq : NULL, S->hits, S->nhits, S->frontier);
}

static u32 prove_mode(const char *s)
{
    if (!strcmp(s, "auto")) return WF_PROVE_AUTO;
    if (!strcmp(s, "off")) return WF_PROVE_OFF;
    if (!strcmp(s, "on")) return WF_PROVE_ON;
    fail("--prove takes auto, off or on, not '%s'", s);
}

static int cmd_scan(int argc, char **argv)
{
    scan_opts o = {0, NULL, NULL, 60, false, true};
    const char *bases = "oeis", *logp = NULL, *v;
    u64 chunk = 10000000000ULL;
    u128 pos[2];
    int npos = 0;
    bool sumq = false;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (opt(argc, argv, &i, "-c", &v)) chunk = num64(v);
        else if (opt(argc, argv, &i, "-b", &v)) bases = v;
        else if (opt(argc, argv, &i, "-n", &v)) {
            const u64 t = num64(v);
            o.near_T = t > (u64)LLONG_MAX ? LLONG_MAX : (long long)t;
        } else if (opt(argc, argv, &i, "-S", &v)) o.state = v;
        else if (opt(argc, argv, &i, "-L", &v)) logp = v;
        else if (opt(argc, argv, &i, "-i", &v)) o.interval = num_in(v, 1, 86400, "interval");
        else if (opt(argc, argv, &i, "--prove", &v)) g_conf.prove = prove_mode(v);
        else if (opt(argc, argv, &i, "--fault-p", &v)) g_conf.fault_p = num64(v);
        else if (!strcmp(argv[i], "-q")) o.quiet = true;
        else if (!strcmp(argv[i], "-Q")) sumq = true;
        else if (argv[i][0] != '-' && npos < 2) pos[npos++] = num128(argv[i]);
        else usage_exit();
    }
    if (npos == 0) usage_exit();
    const u128 start = npos == 2 ? pos[0] : 0, end = pos[npos - 1];
    if (start >= WT_TWO64) fail("START must be < 2^64");
    choose_bases(bases);
    if (logp && !(o.log = fopen(logp, "a"))) fail("cannot open %s: %s", logp, strerror(errno));
    scan_t S = {0};
    const session_t st = run_scan(&S, (u64)start, end, chunk, &o);
    summary(&S, o.near_T, sumq);
    if (!o.quiet) {
        char d[32];
        wt_duration(d, sizeof d, st.seconds);
        fprintf(stderr, "%s this session, %.3g primes/s, %.3g tests/s (GPU kernels %.3f s)\n", d,
                st.seconds > 0 ? (double)st.primes / st.seconds : 0, st.seconds > 0 ? (double)st.tests / st.seconds : 0,
                st.gpu_ms / 1e3);
    }
    if (o.log) fclose(o.log);
    close_gpu();
    const int status = S.flt ? 3 : 0;
    free(S.hits);
    return status;
}

/* ------------------------------------------------------------------ CPU cross-checks */

/* a b mod m (m < 2^64) */
static u64 mul_mod64(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }

static u64 pow_mod64(u64 a, u64 e, u64 m)
{
    u64 r = 1 % m;
    for (a %= m; e; e >>= 1, a = mul_mod64(a, a, m))
        if (e & 1) r = mul_mod64(r, a, m);
    return r;
}

/* Deterministic Miller-Rabin below 2^64 with the seven bases of Jim Sinclair's set, which
 * has no strong pseudoprime below 2^64. */
static bool prime64(u64 n)
{
    static const u64 A[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (n < 2) return false;
    for (u64 d = 2; d < 64; d++)
        if (n % d == 0) return n == d;
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        s++;
    }
    for (size_t i = 0; i < sizeof A / sizeof *A; i++) {
        const u64 a = A[i] % n;
        if (a == 0) continue;
        u64 x = pow_mod64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int k = 1;
        for (; k < s; k++) {
            x = mul_mod64(x, x, n);
            if (x == n - 1) break;
        }
        if (k == s) return false;
    }
    return true;
}

/* a b mod m for any m < 2^128 (a, b < m), most significant bit of b first: r = 2r (+ a),
 * each sum reduced once, with its 129th bit kept by the wrap test. */
static u128 mul_mod128(u128 a, u128 b, u128 m)
{
    u128 r = 0;
    for (int i = 127; i >= 0; i--) {
        const u128 d = r + r;
        r = (d < r || d >= m) ? d - m : d;
        if ((b >> i) & 1) {
            const u128 t = r + a;
            r = (t < r || t >= m) ? t - m : t;
        }
    }
    return r;
}

static u128 pow_mod128(u64 b, u64 e, u128 m)
{
    u128 r = 1 % m, x = b % m;
    for (; e; e >>= 1, x = mul_mod128(x, x, m))
        if (e & 1) r = mul_mod128(r, x, m);
    return r;
}

/* p^-1 mod 2^64 (odd p) by Newton's iteration from the 5-bit inverse p (p p = 1 mod 8). */
static u64 inverse64(u64 p)
{
    u64 x = p;
    for (int k = 0; k < 5; k++) x *= 2 - p * x;
    return x;
}

/* ------------------------------------------------------------------ check, bench */

static int cmd_check(int argc, char **argv)
{
    const char *bases = "known", *v;
    u64 p = 0;
    bool have = false;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (opt(argc, argv, &i, "-b", &v)) bases = v;
        else if (argv[i][0] != '-' && !have) {
            p = num64(argv[i]);
            have = true;
        } else usage_exit();
    }
    if (!have) usage_exit();
    if (p < 3 || !(p & 1)) fail("P must be odd and in [3, 2^64)");
    choose_bases(bases);
    open_gpu();
    const bool prime = prime64(p);
    printf("p = %" PRIu64 " (%s)\n", p, prime ? "prime" : "NOT prime: quotients are meaningless");
    int disagree = 0;
    for (int j = 0; j < g_nbase; j++) {
        const u32 b = g_base[j];
        if (b % p == 0) {
            printf("base %u: p divides b\n", b);
            continue;
        }
        u64 q, rh, rl;
        const int rc = wf_quotients_base(g_gpu, b, &p, &q, &rh, &rl, 1);
        if (rc) fail("wf_quotients_base: %s", wf_strerror(rc));
        /* the residue on the CPU; its floor quotient by p equals q exactly when p is prime */
        const u128 r = pow_mod128(b, p - 1, (u128)p * p);
        const bool ok = (((u128)rh << 64) | rl) == r && (!prime || q == (u64)((r - 1) / p));
        disagree += !ok;
        const long long A = q <= p / 2 ? (long long)q : -(long long)(p - q);
        printf("base %3u: q = %" PRIu64 "  A = %lld%s  [reference %s]\n", b, q, A, q == 0 ? "  <-- WIEFERICH" : "",
               ok ? "agrees" : "DISAGREES");
    }
    close_gpu();
    return disagree ? 1 : 0;
}

/* li(y) - li(x), the expected count of primes in [x, y): the integral of 1 / ln t, taken
 * over u = ln t by Simpson's rule. */
static double primes_between(double x, double y)
{
    const int n = 2000;
    const double a = log(x), h = (log(y) - a) / n;
    double s = 0;
    for (int i = 0; i <= n; i++) {
        const double u = a + i * h, f = exp(u) / u;
        s += (i == 0 || i == n) ? f : (i & 1) ? 4 * f : 2 * f;
    }
    return s * h / 3;
}

static int cmd_bench(int argc, char **argv)
{
    const char *bases = "oeis", *v;
    u64 n0 = 1000000000000000ULL, span = 10000000000ULL;
    bool have = false;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (opt(argc, argv, &i, "-d", &v)) span = num64(v);
        else if (opt(argc, argv, &i, "-b", &v)) bases = v;
        else if (argv[i][0] != '-' && !have) {
            n0 = num64(argv[i]);
            have = true;
        } else usage_exit();
    }
    choose_bases(bases);
    if (n0 + span < n0) fail("N + SPAN does not fit in 64 bits");
    scan_t S = {0};
    const scan_opts o = {0, NULL, NULL, 60, true, true};
    session_t st = run_scan(&S, n0, (u128)n0 + span, span, &o);
    if (st.seconds <= 0) st.seconds = 1e-9;
    const double tps = (double)st.tests / st.seconds;
    printf("bench [%" PRIu64 ", +%" PRIu64 "): %" PRIu64 " primes x %d bases in %.2f s on GPU %d (kernels %.3f s)\n", n0,
           span, st.primes, g_nbase, st.seconds, g_device, st.gpu_ms / 1e3);
    printf("  %.4g primes/s, %.4g tests/s, %.4g numbers/s\n", (double)st.primes / st.seconds, tps,
           (double)st.numbers / st.seconds);
    /* projections: tests of [N, kN) at this rate, and the expected count of new solutions,
     * sum over p of 1/p per base = ln(ln kN / ln N) */
    for (int k = 0; k < 3 && n0 >= 1000 && tps > 0; k++) {
        const double f = k == 0 ? 2 : k == 1 ? 10 : 100, x = (double)n0, y = f * x;
        const double tests = primes_between(x, y) * g_nbase;
        char d[32];
        wt_duration(d, sizeof d, tests / tps);
        printf("  [N, %gN): %.3g tests, ~%s at this rate; heuristic %.3f new solutions over %d bases\n", f, tests, d,
               g_nbase * log(log(y) / log(x)), g_nbase);
    }
    close_gpu();
    free(S.hits);
    return 0;
}

/* ------------------------------------------------------------------ selftest */

/* The base-2 checksums of wieferich_cks.h's three windows (tests/golden/make_golden.py,
 * Python integers: the reference's header has no base-2 column), and its 17-base values
 * (A001220/wieferich_cks.h:2-4) in the order of the "oeis" list. */
static const struct { u64 lo, hi, primes, cks2; u64 oeis[17]; } CK_WINDOWS[] = {
    {1000000000000000ULL, 1000000004000000ULL, 115659ULL, 0x82c9460fec4b2a42ULL,
     {0xf8e9d8adfc5b9653ULL, 0xb8c7b117e7d8d566ULL, 0x00c805e26cadf5a9ULL, 0xb46553b17fef1432ULL, 0x4517810450a108b5ULL,
      0x4c308878022e110bULL, 0x93019cb69192dfdfULL, 0xd98ebbbb270ec8ebULL, 0x65f56c9af042f1ebULL, 0x5d52787b79c28015ULL,
      0xcd08b2f19f6f9e7fULL, 0xe45d09330461b52eULL, 0xeb57d3fa802e2803ULL, 0xe01404527ca951a6ULL, 0xa6a3cbb44312e850ULL,
      0xb44f53280d3aa0f0ULL, 0xf128cb909f21b185ULL}},
    {1000000000000000000ULL, 1000000000002000000ULL, 48427ULL, 0xfa574a8a4dd0bbf7ULL,
     {0x1725b31f655b0226ULL, 0x60a6ff72682a26d2ULL, 0xd8988e0b9135a43cULL, 0xd5969b80a8187475ULL, 0xd00694f6e6354bd0ULL,
      0xd30b154c777dc267ULL, 0x16294824948c2a9dULL, 0x5038a6f0ce44ed96ULL, 0x2f0a1a1d1810fe42ULL, 0x1aff2fd0691a2dafULL,
      0x31e094e7c60a99e8ULL, 0x44152bb55fda44c5ULL, 0x05f992d233d411b1ULL, 0xa62efd0e9552024eULL, 0x29cf767b857da8f3ULL,
      0xcec9ecde3002d809ULL, 0x3b34bf9f7a2b6d93ULL}},
    {9223372036852775808ULL, 9223372036854775808ULL, 45777ULL, 0x50454bd548214c98ULL,
     {0xff626de7fb316232ULL, 0x1d7d16db4165399bULL, 0xdc418aac8bcb404cULL, 0x0debd3cf6b884231ULL, 0x68a32de6eb7b7808ULL,
      0xe9b4fe527ac8558dULL, 0xf53c6b79f162a8cbULL, 0x7cae99135f554f17ULL, 0x2c387a41f6ea09a5ULL, 0x629cbbb17a706ea3ULL,
      0x99871c62d5032924ULL, 0xdbda0f3a9c6c30abULL, 0x5e92dd42e137b077ULL, 0x4ebd44f5dfd65eb9ULL, 0x00c12e131ea2b18cULL,
      0xfd19e5fd0c9df246ULL, 0xf4ac8d4519009025ULL}},
};

/* The known solutions of the "known" bases (the OEIS values the reference's selftest
 * checks, A001220/wieferich.c:699-723), as (base, p) pairs. */
static const struct { u32 b; u64 p; } SOLUTIONS[] = {
    {2, 1093ULL}, {2, 3511ULL}, {3, 11ULL}, {3, 1006003ULL}, {5, 2ULL}, {5, 20771ULL}, {5, 40487ULL},
    {5, 53471161ULL}, {5, 1645333507ULL}, {5, 6692367337ULL}, {5, 188748146801ULL}, {6, 66161ULL}, {6, 534851ULL},
    {6, 3152573ULL}, {7, 5ULL}, {7, 491531ULL}, {10, 3ULL}, {10, 487ULL}, {10, 56598313ULL}, {11, 71ULL},
    {12, 2693ULL}, {12, 123653ULL}, {13, 2ULL}, {13, 863ULL}, {13, 1747591ULL}, {14, 29ULL}, {14, 353ULL},
    {14, 7596952219ULL}, {15, 29131ULL}, {15, 119327070011ULL}, {17, 2ULL}, {17, 3ULL}, {17, 46021ULL},
    {17, 48947ULL}, {17, 478225523351ULL}, {18, 5ULL}, {18, 7ULL}, {18, 37ULL}, {18, 331ULL}, {18, 33923ULL},
    {18, 1284043ULL}, {19, 3ULL}, {19, 7ULL}, {19, 13ULL}, {19, 43ULL}, {19, 137ULL}, {19, 63061489ULL},
    {20, 281ULL}, {20, 46457ULL}, {20, 9377747ULL}, {20, 122959073ULL}, {21, 2ULL}, {22, 13ULL}, {22, 673ULL},
    {22, 1595813ULL}, {22, 492366587ULL}, {22, 9809862296159ULL}, {23, 13ULL}, {23, 2481757ULL}, {23, 13703077ULL},
    {23, 15546404183ULL}, {23, 2549536629329ULL}, {24, 5ULL}, {24, 25633ULL}, {26, 3ULL}, {26, 5ULL}, {26, 71ULL},
    {26, 486999673ULL}, {26, 6695256707ULL}, {28, 3ULL}, {28, 19ULL}, {28, 23ULL}, {29, 2ULL}, {30, 7ULL},
    {30, 160541ULL}, {30, 94727075783ULL}, {30, 303632117562967ULL},
};
#define NSOLUTIONS (sizeof SOLUTIONS / sizeof *SOLUTIONS)

static bool is_solution(u32 b, u64 p)
{
    for (size_t i = 0; i < NSOLUTIONS; i++)
        if (SOLUTIONS[i].b == b && SOLUTIONS[i].p == p) return true;
    return false;
}

static const scan_opts SILENT = {0, NULL, NULL, 60, true, false};

/* Part 0: GPU residues and quotients of random odd numbers of every length (and edge
 * moduli) for five bases, against the CPU's 128-bit arithmetic. */
static int selftest_arith(void)
{
    enum { NQ = 8000 };
    static const u64 edge[] = {3, 5, 7, 1093, 3511, 4294967291ULL, 4294967311ULL, (1ULL << 50) - 27,
                               (1ULL << 50) + 27, (9ULL << 47) - 1, (9ULL << 47) + 1,   /* FP64 class end */
                               (1ULL << 63) - 25, (1ULL << 63) + 29, 18446744073709551557ULL, UINT64_MAX};
    static const u32 tb[] = {2, 3, 30, 1625, 4294967291U};
    u64 *p = malloc(4 * NQ * sizeof *p);
    if (!p) fail("out of memory");
    u64 *q = p + NQ, *rh = q + NQ, *rl = rh + NQ;
    u64 seed = 0x243F6A8885A308D3ULL;      /* splitmix64, the reference selftest's seed */
    size_t n = 0;
    for (; n < sizeof edge / sizeof *edge; n++) p[n] = edge[n];
    while (n < NQ) {
        u64 z = (seed += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        const int bits = 2 + (int)(z % 63);
        const u64 x = (z >> (64 - bits)) | (1ULL << (bits - 1)) | 1;
        if (x >= 3) p[n++] = x;
    }
    int bad = 0;
    for (size_t k = 0; k < sizeof tb / sizeof *tb; k++) {
        const int rc = wf_quotients_base(g_gpu, tb[k], p, q, rh, rl, NQ);
        if (rc) fail("wf_quotients_base: %s", wf_strerror(rc));
        for (int i = 0; i < NQ && bad < 5; i++) {
            const u128 r = pow_mod128(tb[k], p[i] - 1, (u128)p[i] * p[i]);
            if ((((u128)rh[i] << 64) | rl[i]) != r || q[i] != ((u64)r - 1) * inverse64(p[i])) {
                bad++;
                fprintf(stderr, "  quotient mismatch base %u p=%" PRIu64 "\n", tb[k], p[i]);
            }
        }
    }
    fprintf(stderr, "arithmetic: %d GPU residues b^(p-1) mod p^2 for %zu bases vs 128-bit CPU arithmetic: %s\n", NQ,
            sizeof tb / sizeof *tb, bad ? "FAIL" : "ok");
    free(p);
    return bad;
}

/* Part 1: the "known" bases over [0, limit): every known solution below the end and no
 * other, no FLT error, and pi(10^9) when limit = 10^9.  Returns the failures; *end = the
 * scanned end. */
static int selftest_known(u64 limit, u64 *end)
{
    const double t = seconds();
    choose_bases("known");
    scan_t S = {0};
    run_scan(&S, 0, limit, limit / 16 > 1000000 ? limit / 16 : 1000000, &SILENT);
    *end = (u64)S.end;
    int missing = 0, extra = 0;
    for (size_t i = 0; i < S.nhits; i++)
        if (!is_solution(S.hits[i].b, S.hits[i].p)) {
            extra++;
            fprintf(stderr, "  unexpected solution base %u p %" PRIu64 "\n", S.hits[i].b, S.hits[i].p);
        }
    for (size_t i = 0; i < NSOLUTIONS; i++)
        if (SOLUTIONS[i].p < *end && !known_hit(&S, SOLUTIONS[i].b, SOLUTIONS[i].p)) {
            missing++;
            fprintf(stderr, "  missing base %u p %" PRIu64 "\n", SOLUTIONS[i].b, SOLUTIONS[i].p);
        }
    const bool pi_checked = limit == 1000000000ULL, pi_ok = !pi_checked || S.primes == 50847534ULL;
    fprintf(stderr, "scan [0, %" PRIu64 "), %d bases: %" PRIu64 " primes%s, %zu known solutions found, %d missing, %d "
                    "unexpected, %" PRIu64 " FLT errors (%.1f s)\n", *end, g_nbase, S.primes,
            pi_checked ? (pi_ok ? " (= pi(10^9))" : " (WRONG, pi(10^9) = 50847534)") : "", S.nhits - (size_t)extra,
            missing, extra, S.flt, seconds() - t);
    const int bad = missing + extra + (int)S.flt + !pi_ok;
    free(S.hits);
    return bad;
}

/* Part 2: a window of 2e6 numbers around each known solution above the scan, base by base:
 * exactly that solution is found. */
static int selftest_windows(u64 scanned)
{
    const double t = seconds();
    int wins = 0, bad = 0;
    for (size_t i = 0; i < NSOLUTIONS; i++) {
        const u64 p = SOLUTIONS[i].p;
        if (p < scanned) continue;
        char one[16];
        snprintf(one, sizeof one, "%u", SOLUTIONS[i].b);
        choose_bases(one);
        scan_t S = {0};
        run_scan(&S, p > 1000000 ? p - 1000000 : 0, (u128)p + 1000000, 200000, &SILENT);
        if (S.nhits != 1 || S.hits[0].p != p || S.flt) {
            bad++;
            fprintf(stderr, "  window around base %u p %" PRIu64 ": %zu hits\n", SOLUTIONS[i].b, p, S.nhits);
        }
        wins++;
        free(S.hits);
    }
    fprintf(stderr, "windows: %d known solutions above %" PRIu64 " re-found through the scan: %s (%.1f s)\n", wins,
            scanned, bad ? "FAIL" : "ok", seconds() - t);
    return bad;
}

/* Part 3: the checksum windows, the "oeis" bases and base 2. */
static int selftest_checksums(void)
{
    const double t = seconds();
    int bad = 0;
    for (int two = 0; two < 2; two++) {
        choose_bases(two ? "2" : "oeis");
        for (size_t w = 0; w < sizeof CK_WINDOWS / sizeof *CK_WINDOWS; w++) {
            const u64 lo = CK_WINDOWS[w].lo, hi = CK_WINDOWS[w].hi;
            scan_t S = {0};
            run_scan(&S, lo, hi, (hi - lo) / 16, &SILENT);
            bool ok = S.primes == CK_WINDOWS[w].primes && S.flt == 0;
            for (int j = 0; j < g_nbase; j++) {
                const u64 want = two ? CK_WINDOWS[w].cks2 : CK_WINDOWS[w].oeis[j];
                if (S.cks[j] != want) {
                    ok = false;
                    fprintf(stderr, "    base %u: %016" PRIx64 " expected %016" PRIx64 "\n", g_base[j], S.cks[j], want);
                }
            }
            if (!ok) {
                bad++;
                fprintf(stderr, "  checksum window [%" PRIu64 ", %" PRIu64 "), %s: primes %" PRIu64 "\n", lo, hi,
                        two ? "base 2" : "oeis bases", S.primes);
            }
            free(S.hits);
        }
    }
    fprintf(stderr, "checksums: %zu windows x 17 'oeis' bases (wieferich_cks.h) and base 2 (Python integers): %s (%.1f s)\n",
            sizeof CK_WINDOWS / sizeof *CK_WINDOWS, bad ? "FAIL" : "ok", seconds() - t);
    return bad;
}

/* selftest [LIMIT]: the reference's parts (wieferich.c:1821-1900) on the GPU. */
static int cmd_selftest(int argc, char **argv)
{
    u64 limit = 1000000000ULL;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (argv[i][0] != '-') limit = num64(argv[i]);
        else usage_exit();
    }
    choose_bases("2");
    open_gpu();
    u64 scanned = 0;
    int bad = selftest_arith();
    bad += selftest_known(limit, &scanned);
    bad += selftest_windows(scanned);
    bad += selftest_checksums();
    fprintf(stderr, "selftest %s\n", bad ? "FAILED" : "passed");
    close_gpu();
    return bad ? 1 : 0;
}

/* ------------------------------------------------------------------ num, dcheck, stats */

/* num X ...: the GPU's base-2 strong test, strong Lucas test and BPSW verdict per X (the
 * reference's `a350826_cuda num`, A350826/cuda/a350826_cuda.cu:2458-2483). */
static int cmd_num(int argc, char **argv)
{
    u64 *x = malloc((size_t)(argc > 0 ? argc : 1) * sizeof *x);
    uint8_t *f = malloc((size_t)(argc > 0 ? argc : 1));
    if (!x || !f) fail("out of memory");
    int n = 0;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (argv[i][0] != '-') x[n++] = num64(argv[i]);
        else usage_exit();
    }
    if (!n) usage_exit();
    choose_bases("2");
    open_gpu();
    const int rc = wf_bpsw(g_gpu, x, f, (u64)n);
    if (rc) fail("wf_bpsw: %s", wf_strerror(rc));
    for (int i = 0; i < n; i++)
        printf("%" PRIu64 ": gpu sprp2 %d slprp %d prime %d\n", x[i], f[i] & WF_BPSW_SPRP2 ? 1 : 0,
               f[i] & WF_BPSW_LUCAS ? 1 : 0, f[i] & WF_BPSW_PRIME ? 1 : 0);
    close_gpu();
    free(x);
    free(f);
    return 0;
}

/* Whether OUT already has a result line for the chunk [lo, hi). */
static bool already_checked(const char *out, u64 lo, const char *hi)
{
    FILE *f = fopen(out, "r");
    if (!f) return false;
    char *line = NULL, *key = NULL;
    size_t cap = 0;
    bool found = false;
    if (asprintf(&key, " %" PRIu64 " %s ", lo, hi) < 0) fail("out of memory");
    while (!found && getline(&line, &cap, f) >= 0) found = strstr(line, key) != NULL;
    free(line);
    free(key);
    fclose(f);
    return found;
}

/* dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID]: A001220/dcheck.sh's double check on
 * the GPU.  Every STEP-th line (default 16) of a chunk log is recomputed by an engine set up
 * differently from a scan's: prove mode for every piece (small-prime sieve + BPSW instead of
 * the bucket sieve), 16 KiB segments, one ladder per base (a scan derives composite bases
 * from their prime factors), and the 32-bit limb Montgomery ladders (doublings for base 2, the
 * plain binary ladder for any other base) instead of the FP64 p-adic / radix-2^28 window
 * ladders.  The recomputed line must equal the logged one byte
 * for byte: "MATCH lo hi TIME" or "MISMATCH lo hi gpu: LINE check: LINE" is appended to OUT
 * (default CHUNKLOG.dcheck).  Chunks already in OUT are skipped, so a run resumes; with PID,
 * new lines are awaited while that process lives (dcheck.sh:12-45).  THREADS is accepted and
 * ignored.  Exit status 0 when every chunk checked in this run matched, else 1. */
static int cmd_dcheck(int argc, char **argv)
{
    const char *arg[6] = {0};
    int na = 0;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (argv[i][0] != '-' && na < 6) arg[na++] = argv[i];
        else usage_exit();
    }
    if (na < 2) usage_exit();
    const char *logp = arg[0];
    choose_bases(arg[1]);
    const u64 step = na > 2 ? num64(arg[2]) : 16;
    if (step == 0) fail("STEP must be positive");
    if (na > 3) (void)num_in(arg[3], 1, 1024, "threads");
    char *outp = NULL;
    if (na > 4) outp = strdup(arg[4]);
    else if (asprintf(&outp, "%s.dcheck", logp) < 0) outp = NULL;
    if (!outp) fail("out of memory");
    const long pid = na > 5 ? (long)num64(arg[5]) : 0;
    g_conf.prove = WF_PROVE_ON;
    g_conf.seg_bytes = 16384;
    g_conf.derive = WF_DERIVE_OFF;
    g_conf.arith = 1;
    open_gpu();
    FILE *lf = fopen(logp, "r");
    if (!lf) fail("cannot open %s: %s", logp, strerror(errno));
    wf_event *ev = malloc(EVENT_CAP * sizeof *ev);
    if (!ev) fail("out of memory");
    char *line = NULL;
    size_t cap = 0;
    u64 seen = 0, checked = 0, mismatches = 0;
    for (;;) {
        bool progressed = false;
        long mark = ftell(lf);
        ssize_t len;
        while ((len = getline(&line, &cap, lf)) >= 0) {
            if (len == 0 || line[len - 1] != '\n') {    /* the writer is still on this line */
                fseek(lf, mark, SEEK_SET);
                break;
            }
            line[--len] = 0;
            mark = ftell(lf);
            if (++seen % step) continue;
            char los[32], his[32];
            if (sscanf(line, "C %31s %31s", los, his) != 2) fail("%s: not a chunk-log line: %s", logp, line);
            const u64 lo = num64(los);
            const u128 hi = num128(his);
            if (hi <= lo) fail("%s: bad range in: %s", logp, line);
            if (already_checked(outp, lo, his)) continue;
            wf_result r;
            const int rc = wf_scan(g_gpu, lo, hi >= WT_TWO64 ? 0 : (u64)hi, 0, &r, ev, EVENT_CAP, 0, NULL, 0);
            if (rc) fail("chunk [%" PRIu64 ", %s): %s", lo, his, wf_strerror(rc));
            char mine[64 + WF_MAXB * 32];
            if (r.flt) snprintf(mine, sizeof mine, "BAD range [%" PRIu64 ", %s) flt %" PRIu64, lo, his, r.flt);
            else wt_chunk_line(mine, sizeof mine, lo, hi, r.primes, g_base, r.cks, g_nbase);
            FILE *of = fopen(outp, "a");
            if (!of) fail("cannot open %s: %s", outp, strerror(errno));
            if (!strcmp(mine, line)) {
                char ts[32];
                const time_t now = time(NULL);
                strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%S", localtime(&now));
                fprintf(of, "MATCH %" PRIu64 " %s %s\n", lo, his, ts);
            } else {
                fprintf(of, "MISMATCH %" PRIu64 " %s gpu: %s check: %s\n", lo, his, line, mine);
                mismatches++;
            }
            fclose(of);
            checked++;
            progressed = true;
        }
        clearerr(lf);
        const bool writer_alive = pid > 0 && kill((pid_t)pid, 0) == 0;
        if (!writer_alive && !progressed) break;
        if (writer_alive && !progressed) sleep(60);
    }
    fclose(lf);
    free(line);
    free(ev);
    free(outp);
    close_gpu();
    fprintf(stderr, "dcheck: %" PRIu64 " chunks checked, %" PRIu64 " mismatches\n", checked, mismatches);
    return mismatches ? 1 : 0;
}

/* stats [LO] HI: the sieve's statistics for the repository's other prime consumers: twin
 * pairs (A007508), the largest gap, and the run lengths of A158939 (L(p) = the number of
 * strictly increasing gaps after p, A158939/a158939.c:19-24) with the first prime of each. */
static int cmd_stats(int argc, char **argv)
{
    u128 pos[2];
    int npos = 0;
    for (int i = 0; i < argc; i++) {
        if (common_opt(argc, argv, &i)) continue;
        if (argv[i][0] != '-' && npos < 2) pos[npos++] = num128(argv[i]);
        else usage_exit();
    }
    if (npos == 0) usage_exit();
    const u128 lo = npos == 2 ? pos[0] : 0, hi = pos[npos - 1];
    if (hi <= lo) fail("END must be larger than START");
    if (hi > (u128)WF_STATS_MAXHI) fail("END must be <= 2^64 - 2^20 (the gaps after the last primes are needed)");
    choose_bases("2");
    open_gpu();
    wf_stats st;
    const double t0 = seconds();
    const int rc = wf_sieve_stats(g_gpu, (u64)lo, (u64)hi, &st);
    if (rc) fail("wf_sieve_stats: %s", wf_strerror(rc));
    printf("STATS [%" PRIu64 ", %" PRIu64 ") primes %" PRIu64 " twins %" PRIu64 " max_gap %" PRIu64 " after %" PRIu64 "\n",
           (u64)lo, (u64)hi, st.primes, st.twins, st.max_gap, st.max_gap_p);
    for (int n = 1; n < 64; n++)
        if (st.runs[n]) printf("run %2d count %" PRIu64 " first %" PRIu64 "\n", n, st.runs[n], st.first[n]);
    fprintf(stderr, "%.3f s\n", seconds() - t0);
    close_gpu();
    return 0;
}

int main(int argc, char **argv)
{
    static const struct { const char *name; int (*run)(int, char **); } CMDS[] = {
        {"scan", cmd_scan}, {"check", cmd_check}, {"bench", cmd_bench}, {"selftest", cmd_selftest},
        {"num", cmd_num}, {"dcheck", cmd_dcheck}, {"stats", cmd_stats}};
    g_tty = isatty(2);
    if (argc < 2) usage_exit();
    for (size_t i = 0; i < sizeof CMDS / sizeof *CMDS; i++)
        if (!strcmp(argv[1], CMDS[i].name)) return CMDS[i].run(argc - 2, argv + 2);
    usage_exit();
}
