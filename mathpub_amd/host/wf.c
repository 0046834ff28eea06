/*
 * wf.c -- command-line front end of the MI355X Wieferich engine (libwf.so, include/wf.h).
 *
 * A drop-in for A001220/wieferich.c's command line:
 *
 *   wf scan [START] END [-b BASES] [-c CHUNK] [-n T] [-S FILE] [-i SECS] [-L FILE] [-g DEV] [-q] [-Q]
 *   wf check P [-b BASES] [-g DEV]
 *   wf bench [N] [-d SPAN] [-b BASES] [-g DEV]
 *   wf selftest [LIMIT] [-g DEV]
 *   wf num X [X ...] [-g DEV]                     (BPSW tests, as a350826_cuda num)
 *   wf dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID]    (as A001220/dcheck.sh)
 *   wf stats [LO] HI [-g DEV]                     (twin pairs, gaps, runs of increasing gaps)
 *
 * Kept byte for byte from the reference: the number grammar (wieferich.c:300-416), the
 * base-list grammar (:816-849), chunking with END rounded up to a chunk boundary
 * (:1389-1394), the chunk log line "C lo hi primes b:cks" (:1166-1172), the checkpoint
 * file (:1223-1316, resumable by either tool), the FOUND / NEAR / FLT-ERROR / CONFIRMED
 * lines (:1009, :1060, :1071, :1350-1352), SUMMARY (:1505-1523), exit statuses 0 / 1 /
 * 2 / 3 (:271-280, :1535-1544, :1591), the default base lists ("oeis" for scan and bench,
 * "known" for check) and the selftest's three parts (:1821-1900).  Differences: the
 * default CHUNK is 10^10 as in the GPU tool (wieferich_cuda.cu:1591), END may be 2^64 (the
 * reference stops at 2^63), -Q prints the sum of the quotients, and -t/-T are accepted
 * and ignored (no host sieve threads).
 * Chunks are computed in order, so the frontier is the number of completed chunks.
 */
#define _GNU_SOURCE
#include <ctype.h>
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

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

#define MAXHITS 4096
#define GOLD 0x9E3779B97F4A7C15ULL

static bool stderr_tty;
static volatile sig_atomic_t g_stop = 0;
static struct wf_engine *g_eng = NULL;
static int g_dev = 0;

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void die(const char *fmt, ...)
{
    va_list ap;
    if (stderr_tty) fputs("\r\033[K", stderr);
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    exit(1);
}

static const char *fmt_dur(double s, char *buf, size_t n)
{
    if (s < 0 || s != s || s > 1e12) { snprintf(buf, n, "?"); return buf; }
    long t = (long)s;
    if (t >= 86400) snprintf(buf, n, "%ldd%02ldh%02ldm", t / 86400, (t / 3600) % 24, (t / 60) % 60);
    else if (t >= 3600) snprintf(buf, n, "%ldh%02ldm%02lds", t / 3600, (t / 60) % 60, t % 60);
    else if (t >= 60) snprintf(buf, n, "%ldm%02lds", t / 60, t % 60);
    else snprintf(buf, n, "%.1fs", s);
    return buf;
}

/* ---------------------------------------------------------------- numbers */

/* Values are parsed into 128 bits so that 2^64 (the end of the top-of-word range) is
 * representable; anything above 2^64 is rejected. */
#define LIMIT128 (((u128)1) << 64)

static bool mul_ok(u128 a, u128 b, u128 *r)
{
    if (a && b > LIMIT128 / a) return false;
    *r = a * b;
    return *r <= LIMIT128;
}

static bool digits_only(const char *s, size_t n)
{
    if (n == 0) return false;
    for (size_t i = 0; i < n; i++) if (!isdigit((unsigned char)s[i])) return false;
    return true;
}

static bool dec(const char *s, size_t n, u128 *out)
{
    u128 v = 0;
    if (!digits_only(s, n)) return false;
    for (size_t i = 0; i < n; i++) {
        if (!mul_ok(v, 10, &v)) return false;
        v += (u128)(s[i] - '0');
        if (v > LIMIT128) return false;
    }
    *out = v;
    return true;
}

/* One term: decimal, B^E, or MeE with an optional decimal point (wieferich.c:329-378). */
static bool parse_term(const char *s, size_t n, u128 *out)
{
    char buf[64], *c;
    if (n == 0 || n >= sizeof buf) return false;
    memcpy(buf, s, n);
    buf[n] = 0;
    if ((c = strchr(buf, '^'))) {
        *c = 0;
        u128 b, e, r = 1;
        if (!dec(buf, strlen(buf), &b) || !dec(c + 1, strlen(c + 1), &e)) return false;
        if (b <= 1) { *out = e == 0 ? 1 : b; return true; }
        while (e--) if (!mul_ok(r, b, &r)) return false;
        *out = r;
        return true;
    }
    if ((c = strpbrk(buf, "eE"))) {
        *c = 0;
        char *dot = strchr(buf, '.');
        size_t ip = dot ? (size_t)(dot - buf) : strlen(buf);
        const char *frac = dot ? dot + 1 : buf + ip;
        size_t fl = strlen(frac);
        if ((ip && !digits_only(buf, ip)) || (fl && !digits_only(frac, fl))) return false;
        if (ip + fl == 0) return false;
        if (!digits_only(c + 1, strlen(c + 1))) return false;
        while (fl && frac[fl - 1] == '0') fl--;
        errno = 0;
        long e = strtol(c + 1, 0, 10);
        if (errno) return false;
        e -= (long)fl;
        u128 m = 0;
        for (size_t i = 0; i < ip; i++) { if (!mul_ok(m, 10, &m)) return false; m += (u128)(buf[i] - '0'); }
        for (size_t i = 0; i < fl; i++) { if (!mul_ok(m, 10, &m)) return false; m += (u128)(frac[i] - '0'); }
        if (m > LIMIT128) return false;
        if (m == 0) { *out = 0; return true; }
        for (; e > 0; e--) if (!mul_ok(m, 10, &m)) return false;
        for (; e < 0; e++) { if (m % 10) return false; m /= 10; }
        *out = m;
        return true;
    }
    return dec(buf, n, out);
}

/* TERM, TERM+TERM or TERM-TERM (wieferich.c:393-416); result <= 2^64. */
static u128 parse_num128(const char *s)
{
    size_t n = strlen(s);
    for (size_t i = n; i-- > 1;) {
        if (s[i] != '+' && s[i] != '-') continue;
        if (s[i - 1] == 'e' || s[i - 1] == 'E') continue;
        u128 a, b;
        if (!parse_term(s, i, &a) || !parse_term(s + i + 1, n - i - 1, &b)) die("bad number '%s'", s);
        if (s[i] == '+') {
            if (a + b > LIMIT128) die("overflow in '%s'", s);
            return a + b;
        }
        if (b > a) die("negative result in '%s'", s);
        return a - b;
    }
    u128 v;
    if (!parse_term(s, n, &v)) die("bad number '%s'", s);
    return v;
}

static u64 parse_num(const char *s)
{
    u128 v = parse_num128(s);
    if (v >= LIMIT128) die("overflow in '%s'", s);
    return (u64)v;
}

static int parse_int(const char *s, int lo, int hi, const char *what)
{
    u64 v = parse_num(s);
    if (v < (u64)lo || v > (u64)hi) die("%s must be between %d and %d", what, lo, hi);
    return (int)v;
}

static inline u64 mix64(u64 z)
{
    z += GOLD;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* ---------------------------------------------------------------- bases */

static const u32 OEIS_BASES[] = {3, 5, 6, 7, 10, 12, 13, 14, 15, 17, 18, 19, 20, 22, 23, 26, 30};
static const u32 KNOWN_BASES[] = {2, 3, 5, 6, 7, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 29, 30};
static int g_nb = 0;
static u32 g_b[WF_MAXB];

static int cmp_u32(const void *a, const void *b)
{
    u32 x = *(const u32 *)a, y = *(const u32 *)b;
    return x < y ? -1 : x > y;
}

static void push_base(u32 *list, int *n, u32 b)
{
    for (int i = 0; i < *n; i++) if (list[i] == b) return;
    if (*n >= WF_MAXB) die("too many bases (max %d)", WF_MAXB);
    list[(*n)++] = b;
}

/* Comma list of numbers, LO-HI ranges, "oeis", "known" (wieferich.c:816-849). */
static void parse_bases(const char *s)
{
    u32 list[WF_MAXB];
    int n = 0;
    char buf[1024];
    if (strlen(s) >= sizeof buf) die("base list too long");
    strcpy(buf, s);
    for (char *tok = strtok(buf, ","); tok; tok = strtok(NULL, ",")) {
        if (!strcmp(tok, "oeis")) {
            for (size_t i = 0; i < sizeof OEIS_BASES / sizeof *OEIS_BASES; i++) push_base(list, &n, OEIS_BASES[i]);
            continue;
        }
        if (!strcmp(tok, "known")) {
            for (size_t i = 0; i < sizeof KNOWN_BASES / sizeof *KNOWN_BASES; i++) push_base(list, &n, KNOWN_BASES[i]);
            continue;
        }
        char *dash = strchr(tok, '-');
        u64 lo, hi;
        if (dash) { *dash = 0; lo = parse_num(tok); hi = parse_num(dash + 1); }
        else lo = hi = parse_num(tok);
        if (lo < 2 || hi < lo || hi >= (1ULL << 32)) die("bad base or range '%s'", tok);
        for (u64 b = lo; b <= hi; b++) push_base(list, &n, (u32)b);
    }
    if (n == 0) die("between 1 and %d bases are supported", WF_MAXB);
    qsort(list, (size_t)n, sizeof *list, cmp_u32);
    g_nb = n;
    memcpy(g_b, list, (size_t)n * sizeof *list);
}

static const char *anum_for(u32 b)
{
    static const char *const A[] = {"A014127", "A123692", "A212583", "A123693", "A045616", "A111027", "A128667",
                                    "A234810", "A242741", "A128668", "A244260", "A090968", "A242982", "A298951",
                                    "A128669", "A306255", "A306256"};
    if (b == 2) return "A001220";
    for (size_t i = 0; i < sizeof OEIS_BASES / sizeof *OEIS_BASES; i++) if (OEIS_BASES[i] == b) return A[i];
    return NULL;
}

static void bases_str(char *buf, size_t n)
{
    size_t k = 0;
    buf[0] = 0;
    for (int j = 0; j < g_nb && k < n; j++) {
        int w = snprintf(buf + k, n - k, "%s%u", j ? "," : "", g_b[j]);
        if (w < 0) break;
        k += (size_t)w;
    }
}

/* ---------------------------------------------------------------- engine */

static wf_config g_cfg;   /* zero fields: engine defaults */

/* --prove auto|off|on: prove mode of the engine (wf.h, WF_PROVE_*) */
static u32 parse_prove(const char *s)
{
    if (!strcmp(s, "auto")) return WF_PROVE_AUTO;
    if (!strcmp(s, "off")) return WF_PROVE_OFF;
    if (!strcmp(s, "on")) return WF_PROVE_ON;
    die("--prove takes auto, off or on, not '%s'", s);
    return 0;
}

static void engine_open(void)
{
    if (g_eng) return;
    int rc = wf_open(g_dev, &g_cfg, &g_eng);
    if (rc) die("cannot open GPU %d: %s", g_dev, wf_strerror(rc));
    rc = wf_set_bases(g_eng, g_b, g_nb);
    if (rc) {
        char b[1024];
        bases_str(b, sizeof b);
        die("bases %s: %s", b, wf_strerror(rc));
    }
}

/* ---------------------------------------------------------------- scan state */

typedef struct { u32 b; u64 p; u64 chunk; bool announced; } hit_t;

static struct {
    u64 start, chunk, nchunks;
    u128 end;                       /* may be 2^64 */
    u64 frontier;
    u64 primes, tests, near, flt;
    u64 cks[WF_MAXB], sumq[WF_MAXB];
    int nhits;
    hit_t hits[MAXHITS];
} S;

static long long g_near = 0;
static bool g_print = true;
static FILE *g_log = NULL;

static u64 chunk_lo(u64 i) { return S.start + i * S.chunk; }
/* end of chunk i as 128 bits (2^64 possible) */
static u128 chunk_hi(u64 i)
{
    u128 h = (u128)S.start + (u128)(i + 1) * S.chunk;
    return h < S.end ? h : S.end;
}

static void print_u128(FILE *f, u128 v)
{
    if (v >> 64) fprintf(f, "18446744073709551616");   /* only 2^64 can occur */
    else fprintf(f, "%" PRIu64, (u64)v);
}

/* The chunk-log line of [lo, hi) without its newline: "C lo hi primes b:cks ..."
 * (wieferich.c:1166-1172). */
static void format_chunk_line(char *buf, size_t n, u64 lo, u128 hi, u64 primes, const u64 *cks)
{
    size_t k = (size_t)snprintf(buf, n, "C %" PRIu64 " ", lo);
    if (hi >> 64) k += (size_t)snprintf(buf + k, n - k, "18446744073709551616");
    else k += (size_t)snprintf(buf + k, n - k, "%" PRIu64, (u64)hi);
    k += (size_t)snprintf(buf + k, n - k, " %" PRIu64, primes);
    for (int j = 0; j < g_nb && k < n; j++) k += (size_t)snprintf(buf + k, n - k, " %u:%016" PRIx64, g_b[j], cks[j]);
}

static void record_hit(u32 b, u64 p, u64 idx)
{
    static bool warned = false;
    bool dup = false;
    for (int i = 0; i < S.nhits; i++) if (S.hits[i].b == b && S.hits[i].p == p) dup = true;
    if (!dup && S.nhits < MAXHITS) {
        hit_t *h = &S.hits[S.nhits++];
        h->b = b; h->p = p; h->chunk = idx; h->announced = false;
    } else if (!dup && !warned) {
        warned = true;
        fprintf(stderr, "\nwarning: more than %d solutions; later ones are printed but not kept\n", MAXHITS);
    }
    if (g_print && !dup) {
        if (stderr_tty) fputs("\r\033[K", stderr);
        const char *a = anum_for(b);
        printf("FOUND base %u p %" PRIu64 "%s%s%s\n", b, p, a ? " (" : "", a ? a : "", a ? ")" : "");
        fflush(stdout);
    }
}

static int announce(void)
{
    int n = 0;
    for (int i = 0; i < S.nhits; i++) {
        hit_t *h = &S.hits[i];
        if (h->chunk >= S.frontier) continue;
        n++;
        if (!h->announced) {
            h->announced = true;
            if (g_print) {
                if (stderr_tty) fputs("\r\033[K", stderr);
                const char *a = anum_for(h->b);
                printf("CONFIRMED base %u p %" PRIu64 "%s%s%s (all primes in [%" PRIu64 ", ", h->b, h->p,
                       a ? " (" : "", a ? a : "", a ? ")" : "", S.start);
                print_u128(stdout, (u128)S.start + (u128)S.frontier * S.chunk);
                printf(") tested)\n");
                fflush(stdout);
            }
        }
    }
    return n;
}

/* ---- checkpoint (format of wieferich.c:1223-1316) ---- */

static void save_state(const char *path)
{
    char tmp[PATH_MAX];
    snprintf(tmp, sizeof tmp, "%s.tmp", path);
    FILE *f = fopen(tmp, "w");
    if (!f) { fprintf(stderr, "\nwarning: cannot write %s: %s\n", tmp, strerror(errno)); return; }
    char b[1024];
    bases_str(b, sizeof b);
    fprintf(f, "wieferich checkpoint 1\nbases %s\nstart %" PRIu64 "\nchunk %" PRIu64 "\nend ", b, S.start, S.chunk);
    print_u128(f, S.end);
    fprintf(f, "\nfrontier %" PRIu64 "\nprimes %" PRIu64 "\ntests %" PRIu64 "\nnear %" PRIu64 "\nflt %" PRIu64 "\n",
            S.frontier, S.primes, S.tests, S.near, S.flt);
    for (int j = 0; j < g_nb; j++) fprintf(f, "cks %u %016" PRIx64 "\n", g_b[j], S.cks[j]);
    for (int i = 0; i < S.nhits; i++)
        if (S.hits[i].chunk < S.frontier)
            fprintf(f, "hit %u %" PRIu64 " %" PRIu64 "\n", S.hits[i].b, S.hits[i].p, S.hits[i].chunk);
    fputs("end\n", f);
    bool bad = ferror(f) != 0;
    if (fclose(f) != 0) bad = true;
    if (bad || rename(tmp, path) != 0) {
        fprintf(stderr, "\nwarning: cannot update %s: %s (previous checkpoint kept)\n", path, strerror(errno));
        remove(tmp);
    }
}

static bool load_state(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        if (errno == ENOENT) return false;
        die("cannot read %s: %s", path, strerror(errno));
    }
    char line[4096], bases[4096] = "";
    if (!fgets(line, sizeof line, f) || strncmp(line, "wieferich checkpoint 1", 22) != 0)
        die("%s is not a wieferich checkpoint", path);
    u64 start = 0, chunk = 0, frontier = 0, endv = 0;
    bool complete = false, seen[WF_MAXB] = {false};
    while (fgets(line, sizeof line, f)) {
        if (sscanf(line, "bases %4000s", bases) == 1) continue;
        if (sscanf(line, "start %" SCNu64, &start) == 1) continue;
        if (sscanf(line, "chunk %" SCNu64, &chunk) == 1) continue;
        if (sscanf(line, "end %" SCNu64, &endv) == 1) continue;
        if (sscanf(line, "frontier %" SCNu64, &frontier) == 1) continue;
        if (sscanf(line, "primes %" SCNu64, &S.primes) == 1) continue;
        if (sscanf(line, "tests %" SCNu64, &S.tests) == 1) continue;
        if (sscanf(line, "near %" SCNu64, &S.near) == 1) continue;
        if (sscanf(line, "flt %" SCNu64, &S.flt) == 1) continue;
        if (!strncmp(line, "cks ", 4)) {
            u32 b;
            u64 v;
            if (sscanf(line, "cks %u %" SCNx64, &b, &v) != 2) die("%s: bad cks line", path);
            int j = -1;
            for (int k = 0; k < g_nb; k++) if (g_b[k] == b) j = k;
            if (j < 0) die("%s: base %u is not in the current base list", path, b);
            S.cks[j] = v;
            seen[j] = true;
            continue;
        }
        if (!strncmp(line, "hit ", 4)) {
            hit_t h = {0, 0, 0, false};
            if (sscanf(line, "hit %u %" SCNu64 " %" SCNu64, &h.b, &h.p, &h.chunk) != 3) die("%s: bad hit line", path);
            if (S.nhits < MAXHITS) S.hits[S.nhits++] = h;
            continue;
        }
        if (!strncmp(line, "end", 3)) { complete = true; break; }
    }
    fclose(f);
    if (!complete) die("%s is truncated", path);
    char cur[4096];
    bases_str(cur, sizeof cur);
    if (strcmp(cur, bases) != 0) die("%s was written for bases %s, not %s", path, bases, cur);
    for (int j = 0; j < g_nb; j++)
        if (!seen[j]) die("%s has no checksum for base %u", path, g_b[j]);
    if (start != S.start || chunk != S.chunk)
        die("%s was written for start=%" PRIu64 " chunk=%" PRIu64 ", not start=%" PRIu64 " chunk=%" PRIu64, path, start,
            chunk, S.start, S.chunk);
    if (frontier > S.nchunks)
        die("%s is past the new END (frontier %" PRIu64 " > %" PRIu64 " chunks)", path, frontier, S.nchunks);
    (void)endv;
    S.frontier = frontier;
    return true;
}

static void on_sigint(int sig)
{
    (void)sig;
    g_stop = 1;
}

typedef struct { double seconds, gpu_ms; u64 primes, tests; u128 numbers; } scan_stats_t;

/* Fold one finished chunk (scan_chunk's fold, wieferich.c:1156-1176) and print its events. */
static void fold_chunk(u64 idx, const wf_result *r, const wf_event *ev, u64 nev)
{
    for (u64 k = 0; k < nev; k++) {
        const wf_event *e = &ev[k];
        const u32 b = g_b[e->j];
        if (e->type == WF_EV_FOUND) {
            record_hit(b, e->p, idx);
        } else if (e->type == WF_EV_NEAR) {
            if (g_print) {
                const long long A = e->q <= e->p / 2 ? (long long)e->q : -(long long)(e->p - e->q);
                if (stderr_tty) fputs("\r\033[K", stderr);
                printf("NEAR base %u p %" PRIu64 " A %lld\n", b, e->p, A);
            }
        } else {
            if (stderr_tty) fputs("\r\033[K", stderr);
            printf("FLT-ERROR base %u p %" PRIu64 " r = %016" PRIx64 "%016" PRIx64 "\n", b, e->p, e->r_hi, e->r_lo);
        }
    }
    fflush(stdout);
    S.primes += r->primes;
    S.tests += r->tests;
    S.near += r->near;
    S.flt += r->flt;
    for (int j = 0; j < g_nb; j++) { S.cks[j] += r->cks[j]; S.sumq[j] += r->sumq[j]; }
    if (g_log) {
        char line[64 + WF_MAXB * 32];
        format_chunk_line(line, sizeof line, chunk_lo(idx), chunk_hi(idx), r->primes, r->cks);
        fprintf(g_log, "%s\n", line);
        fflush(g_log);
    }
    S.frontier++;
}

/* Scan [start, end) in chunks, in order (run_scan, wieferich.c:1382-1495). */
static scan_stats_t run_scan(u64 start, u128 end, u64 chunk, const char *state, int interval, bool quiet)
{
    if (end <= start) die("END must be larger than START");
    if (chunk == 0) die("CHUNK must be positive");
    S.start = start;
    S.chunk = chunk;
    const u128 n = (end - start + chunk - 1) / chunk;
    S.nchunks = (u64)n;
    S.end = (u128)start + n * chunk;
    if (S.end > LIMIT128) die("END must be <= 2^64 (after rounding up to the chunk size)");
    if (S.end != end && !quiet) {
        fprintf(stderr, "note: END rounded up to the chunk boundary ");
        print_u128(stderr, S.end);
        fputc('\n', stderr);
    }
    S.frontier = 0;
    S.primes = S.tests = S.near = S.flt = 0;
    memset(S.cks, 0, sizeof S.cks);
    memset(S.sumq, 0, sizeof S.sumq);
    S.nhits = 0;
    if (state && load_state(state) && !quiet)
        fprintf(stderr, "resuming from %s: %" PRIu64 " of %" PRIu64 " chunks done (position %" PRIu64 ", %" PRIu64
                        " primes tested)\n", state, S.frontier, S.nchunks, chunk_lo(S.frontier), S.primes);
    const u64 frontier0 = S.frontier, primes0 = S.primes, tests0 = S.tests;
    engine_open();
    if (!quiet) {
        char info[256], b[1024];
        wf_info(g_eng, info, sizeof info);
        bases_str(b, sizeof b);
        fprintf(stderr, "GPU %d: %s\nscanning [%" PRIu64 ", ", g_dev, info, S.start);
        print_u128(stderr, S.end);
        fprintf(stderr, ") in %" PRIu64 " chunks of %" PRIu64 ", bases %s\n", S.nchunks, S.chunk, b);
    }
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_handler = on_sigint;
    sigaction(SIGINT, &sa, NULL);
    sigaction(SIGTERM, &sa, NULL);

    const u64 EVCAP = 1 << 16;
    wf_event *ev = malloc(EVCAP * sizeof *ev);
    if (!ev) die("out of memory");
    const double t0 = now();
    double last_status = t0, last_ckpt = t0, gpu_ms = 0;
    char b1[32], b2[32];
    while (S.frontier < S.nchunks && !g_stop) {
        const u64 idx = S.frontier;
        const u64 lo = chunk_lo(idx);
        const u128 hi128 = chunk_hi(idx);
        const u64 hi = hi128 >= LIMIT128 ? 0 : (u64)hi128;
        wf_result r;
        int rc = wf_scan(g_eng, lo, hi, g_near, &r, ev, EVCAP, 0, NULL, 0);
        if (rc == WF_E_OVERFLOW && r.nev > EVCAP)
            die("chunk [%" PRIu64 ", ...): %" PRIu64 " events exceed the event buffer (%" PRIu64 "); use a smaller -n or CHUNK",
                lo, r.nev, EVCAP);
        if (rc) die("chunk starting at %" PRIu64 ": %s", lo, wf_strerror(rc));
        wf_timing tm;
        wf_last_timing(g_eng, &tm);
        gpu_ms += tm.total_ms;
        fold_chunk(idx, &r, ev, r.nev < EVCAP ? r.nev : EVCAP);
        const int nconf = announce();
        const double t = now();
        const bool show = stderr_tty ? (t - last_status >= 1.0) : (t - last_status >= interval);
        if (show && !quiet) {
            last_status = t;
            const double done = (double)(S.frontier - frontier0) * (double)S.chunk;
            const double rate = done / (t - t0), trate = (double)(S.tests - tests0) / (t - t0);
            const double eta = rate > 0 ? (double)(S.nchunks - S.frontier) * (double)S.chunk / rate : -1;
            fprintf(stderr, "%s%s  pos %.6g (%.2f%%)  %.3g/s  %.3g tests/s  %.4g primes  ETA %s  found %d (%d confirmed)",
                    stderr_tty ? "\r\033[K" : "progress: ", fmt_dur(t - t0, b1, sizeof b1),
                    (double)chunk_lo(S.frontier), 100.0 * (double)S.frontier / (double)S.nchunks, rate, trate,
                    (double)S.primes, fmt_dur(eta, b2, sizeof b2), S.nhits, nconf);
            if (g_near > 0) fprintf(stderr, "  near %" PRIu64, S.near);
            if (S.flt) fprintf(stderr, "  FLT-ERRORS %" PRIu64, S.flt);
            if (!stderr_tty) fputc('\n', stderr);
            fflush(stderr);
        }
        if (state && t - last_ckpt >= interval) {
            last_ckpt = t;
            save_state(state);
        }
    }
    const double t1 = now();
    if (stderr_tty && !quiet) fputs("\r\033[K", stderr);
    announce();
    if (state) save_state(state);
    if (g_stop && S.frontier < S.nchunks && !quiet)
        fprintf(stderr, "stopped at position %" PRIu64 " (%" PRIu64 " of %" PRIu64 " chunks complete)%s\n",
                chunk_lo(S.frontier), S.frontier, S.nchunks, state ? ", checkpoint saved" : "");
    free(ev);
    scan_stats_t st;
    st.seconds = t1 - t0;
    st.gpu_ms = gpu_ms;
    st.primes = S.primes - primes0;
    st.tests = S.tests - tests0;
    st.numbers = (u128)(S.frontier - frontier0) * S.chunk;
    return st;
}

static void print_summary(bool with_sumq)
{
    u128 pos = (u128)S.start + (u128)S.frontier * S.chunk;
    if (pos > S.end) pos = S.end;
    printf("SUMMARY [%" PRIu64 ", ", S.start);
    print_u128(stdout, pos);
    printf(") primes %" PRIu64 " tests %" PRIu64 " flt_errors %" PRIu64, S.primes, S.tests, S.flt);
    if (g_near > 0) printf(" near(|A|<=%lld) %" PRIu64, g_near, S.near);
    putchar('\n');
    for (int j = 0; j < g_nb; j++) {
        const char *a = anum_for(g_b[j]);
        printf("base %3u %-8s cks %016" PRIx64 "  solutions:", g_b[j], a ? a : "", S.cks[j]);
        int n = 0;
        for (int i = 0; i < S.nhits; i++)
            if (S.hits[i].b == g_b[j] && S.hits[i].chunk < S.frontier) { printf(" %" PRIu64, S.hits[i].p); n++; }
        if (!n) printf(" none");
        putchar('\n');
        if (with_sumq) printf("sumq base %u %016" PRIx64 "\n", g_b[j], S.sumq[j]);
    }
    fflush(stdout);
}

/* ---------------------------------------------------------------- commands */

static void usage(void)
{
    fputs("usage: wf scan [START] END [-b BASES] [-c CHUNK] [-n T] [-S FILE] [-i SECS] [-L FILE] [-g DEV] [-q] [-Q]\n"
          "       wf check P [-b BASES] [-g DEV]\n"
          "       wf bench [N] [-d SPAN] [-b BASES] [-g DEV]\n"
          "       wf selftest [LIMIT] [-g DEV]\n"
          "       wf num X [X ...] [-g DEV]          BPSW tests of single numbers on the GPU\n"
          "       wf dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID] [-g DEV]   recompute every STEP-th chunk\n"
          "       wf stats [LO] HI [-g DEV]          primes, twin pairs, largest gap, runs of increasing gaps\n"
          "scan --prove auto|off|on: sieve short chunks with small primes only and prove the rest by BPSW\n"
          "BASES: list/ranges like 3,5,7 or 2-30; 'oeis' (default) = 3,5,6,7,10,12,13,14,15,17,18,19,20,22,23,26,30;\n"
          "       'known' = the bases 2..30 that are not perfect powers (default of check)\n"
          "END may be 2^64\n", stderr);
    exit(2);
}

static int cmd_scan(int argc, char **argv)
{
    u128 pos[2];
    int npos = 0, interval = 60;
    u64 chunk = 10000000000ULL;
    const char *state = NULL, *logf = NULL, *bases = "oeis";
    bool quiet = false, sumq = false;
    for (int i = 0; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "-c") && i + 1 < argc) chunk = parse_num(argv[++i]);
        else if (!strcmp(a, "-b") && i + 1 < argc) bases = argv[++i];
        else if (!strcmp(a, "-n") && i + 1 < argc) {
            u64 v = parse_num(argv[++i]);
            g_near = v > (u64)LLONG_MAX ? LLONG_MAX : (long long)v;
        }
        else if (!strcmp(a, "-S") && i + 1 < argc) state = argv[++i];
        else if (!strcmp(a, "-L") && i + 1 < argc) logf = argv[++i];
        else if (!strcmp(a, "-i") && i + 1 < argc) interval = parse_int(argv[++i], 1, 86400, "interval");
        else if ((!strcmp(a, "-t") || !strcmp(a, "-T")) && i + 1 < argc) (void)parse_int(argv[++i], 1, 1024, "threads");
        else if (!strcmp(a, "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (!strcmp(a, "-q")) quiet = true;
        else if (!strcmp(a, "-Q")) sumq = true;
        else if (!strcmp(a, "--prove") && i + 1 < argc) g_cfg.prove = parse_prove(argv[++i]);
        else if (a[0] != '-' && npos < 2) pos[npos++] = parse_num128(a);
        else usage();
    }
    if (npos == 0) usage();
    const u128 start = npos == 2 ? pos[0] : 0, end = npos == 2 ? pos[1] : pos[0];
    if (start >= LIMIT128) die("START must be < 2^64");
    parse_bases(bases);
    if (logf && !(g_log = fopen(logf, "a"))) die("cannot open %s: %s", logf, strerror(errno));
    scan_stats_t st = run_scan((u64)start, end, chunk, state, interval, quiet);
    print_summary(sumq);
    if (!quiet) {
        char b1[32];
        fprintf(stderr, "%s this session, %.3g primes/s, %.3g tests/s (GPU kernels %.3f s)\n",
                fmt_dur(st.seconds, b1, sizeof b1), st.seconds > 0 ? (double)st.primes / st.seconds : 0,
                st.seconds > 0 ? (double)st.tests / st.seconds : 0, st.gpu_ms / 1e3);
    }
    if (g_log) fclose(g_log);
    if (g_eng) wf_close(g_eng);
    return S.flt ? 3 : 0;
}

/* ---- independent checks for check / selftest ---- */

static u64 mulmod64(u64 a, u64 b, u64 m) { return (u64)(((u128)a * b) % m); }

static u64 powmod64(u64 a, u64 e, u64 m)
{
    u64 r = 1 % m;
    a %= m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod64(r, a, m);
        a = mulmod64(a, a, m);
    }
    return r;
}

/* deterministic Miller-Rabin below 2^64 (the 12 prime bases <= 37) */
static bool is_prime64(u64 n)
{
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (int i = 0; i < 12; i++) {
        if (n == small[i]) return true;
        if (n % small[i] == 0) return false;
    }
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++) {
        u64 x = powmod64(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s && comp; r++) {
            x = mulmod64(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

/* shift-and-add a b mod m for any m < 2^128 (the carry of 2a and r + a is kept) */
static u128 addmod(u128 a, u128 b, u128 m)
{
    u128 s = a + b;
    if (s < a || s >= m) s -= m;
    return s;
}

static u128 mulmod_ref(u128 a, u128 b, u128 m)
{
    u128 r = 0;
    a %= m;
    b %= m;
    while (b) {
        if (b & 1) r = addmod(r, a, m);
        a = addmod(a, a, m);
        b >>= 1;
    }
    return r;
}

static u128 powmod_ref(u64 b, u64 e, u128 m)
{
    u128 r = 1 % m, x = b % m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod_ref(r, x, m);
        x = mulmod_ref(x, x, m);
    }
    return r;
}

static int cmd_check(int argc, char **argv)
{
    u64 p = 0;
    const char *bases = "known";
    bool havep = false;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "-b") && i + 1 < argc) bases = argv[++i];
        else if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-' && !havep) { p = parse_num(argv[i]); havep = true; }
        else usage();
    }
    if (!havep) usage();
    if (p < 3 || !(p & 1)) die("P must be odd and in [3, 2^64)");
    parse_bases(bases);
    engine_open();
    const bool prime = is_prime64(p);
    printf("p = %" PRIu64 " (%s)\n", p, prime ? "prime" : "NOT prime: quotients are meaningless");
    int bad = 0;
    for (int j = 0; j < g_nb; j++) {
        const u32 b = g_b[j];
        if (b % p == 0) { printf("base %u: p divides b\n", b); continue; }
        u64 q, rh, rl;
        int rc = wf_quotients_base(g_eng, b, &p, &q, &rh, &rl, 1);
        if (rc) die("wf_quotients_base: %s", wf_strerror(rc));
        const u128 rf = ((u128)rh << 64) | rl;
        const u128 rr = powmod_ref(b, p - 1, (u128)p * p);
        const u64 qr = (u64)((rr - 1) / p);
        const long long A = q <= p / 2 ? (long long)q : -(long long)(p - q);
        const bool ok = rf == rr && (!prime || q == qr);
        if (!ok) bad++;
        printf("base %3u: q = %" PRIu64 "  A = %lld%s  [reference %s]\n", b, q, A, q == 0 ? "  <-- WIEFERICH" : "",
               ok ? "agrees" : "DISAGREES");
    }
    wf_close(g_eng);
    return bad ? 1 : 0;
}

static double li_diff(double x, double y)
{
    const int n = 2000;
    double lx = log(x), ly = log(y), h = (ly - lx) / n, s = 0;
    for (int i = 0; i < n; i++) {
        double u = lx + (i + 0.5) * h;
        s += exp(u) / u * h;
    }
    return s;
}

static int cmd_bench(int argc, char **argv)
{
    u64 n0 = 1000000000000000ULL, span = 10000000000ULL;
    const char *bases = "oeis";
    bool haven = false;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "-d") && i + 1 < argc) span = parse_num(argv[++i]);
        else if (!strcmp(argv[i], "-b") && i + 1 < argc) bases = argv[++i];
        else if ((!strcmp(argv[i], "-t") || !strcmp(argv[i], "-T")) && i + 1 < argc) (void)parse_int(argv[++i], 1, 1024, "threads");
        else if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-' && !haven) { n0 = parse_num(argv[i]); haven = true; }
        else usage();
    }
    parse_bases(bases);
    g_print = true;
    if (n0 + span < n0) die("N + SPAN does not fit in 64 bits");
    u64 chunk = span;
    scan_stats_t st = run_scan(n0, (u128)n0 + span, chunk, NULL, 60, true);
    if (st.seconds <= 0) st.seconds = 1e-9;
    const double tps = (double)st.tests / st.seconds, pps = (double)st.primes / st.seconds;
    printf("bench [%" PRIu64 ", +%" PRIu64 "): %" PRIu64 " primes x %d bases in %.2f s on GPU %d (kernels %.3f s)\n", n0,
           span, st.primes, g_nb, st.seconds, g_dev, st.gpu_ms / 1e3);
    printf("  %.4g primes/s, %.4g tests/s, %.4g numbers/s\n", pps, tps, (double)st.numbers / st.seconds);
    if (n0 >= 1000 && tps > 0) {
        const double targets[] = {2, 10, 100};
        for (int i = 0; i < 3; i++) {
            double x = (double)n0, y = x * targets[i];
            double tests = li_diff(x, y) * g_nb;
            double exp_new = g_nb * log(log(y) / log(x));
            char b1[32];
            printf("  [N, %gN): %.3g tests, ~%s at this rate; heuristic %.3f new solutions over %d bases\n", targets[i],
                   tests, fmt_dur(tests / tps, b1, sizeof b1), exp_new, g_nb);
        }
    }
    wf_close(g_eng);
    return 0;
}

/* ---- selftest ---- */

static u64 rng_state = 0x243F6A8885A308D3ULL;
static u64 rnd(void) { return mix64(rng_state += GOLD); }

/* base-2 checksums of the three checksum windows of wieferich_cks.h, from Python big
 * integers (tests/golden/make_golden.py); the reference's header has no base-2 column */
static const struct { u64 lo, hi, primes, cks; } CKWIN2[] = {
    {1000000000000000ULL, 1000000004000000ULL, 115659ULL, 0x82c9460fec4b2a42ULL},
    {1000000000000000000ULL, 1000000000002000000ULL, 48427ULL, 0xfa574a8a4dd0bbf7ULL},
    {9223372036852775808ULL, 9223372036854775808ULL, 45777ULL, 0x50454bd548214c98ULL},
};

/* The checksum windows of A001220/wieferich_cks.h:2-4 for the "oeis" bases (17 values per
 * window, in OEIS_BASES order), the reference's own cross-check data. */
static const struct { u64 lo, hi, primes; u64 cks[17]; } CKWIN_OEIS[] = {
    {1000000000000000ULL, 1000000004000000ULL, 115659ULL,
     {0xf8e9d8adfc5b9653ULL, 0xb8c7b117e7d8d566ULL, 0x00c805e26cadf5a9ULL, 0xb46553b17fef1432ULL, 0x4517810450a108b5ULL, 0x4c308878022e110bULL, 0x93019cb69192dfdfULL, 0xd98ebbbb270ec8ebULL, 0x65f56c9af042f1ebULL, 0x5d52787b79c28015ULL, 0xcd08b2f19f6f9e7fULL, 0xe45d09330461b52eULL, 0xeb57d3fa802e2803ULL, 0xe01404527ca951a6ULL, 0xa6a3cbb44312e850ULL, 0xb44f53280d3aa0f0ULL, 0xf128cb909f21b185ULL}},
    {1000000000000000000ULL, 1000000000002000000ULL, 48427ULL,
     {0x1725b31f655b0226ULL, 0x60a6ff72682a26d2ULL, 0xd8988e0b9135a43cULL, 0xd5969b80a8187475ULL, 0xd00694f6e6354bd0ULL, 0xd30b154c777dc267ULL, 0x16294824948c2a9dULL, 0x5038a6f0ce44ed96ULL, 0x2f0a1a1d1810fe42ULL, 0x1aff2fd0691a2dafULL, 0x31e094e7c60a99e8ULL, 0x44152bb55fda44c5ULL, 0x05f992d233d411b1ULL, 0xa62efd0e9552024eULL, 0x29cf767b857da8f3ULL, 0xcec9ecde3002d809ULL, 0x3b34bf9f7a2b6d93ULL}},
    {9223372036852775808ULL, 9223372036854775808ULL, 45777ULL,
     {0xff626de7fb316232ULL, 0x1d7d16db4165399bULL, 0xdc418aac8bcb404cULL, 0x0debd3cf6b884231ULL, 0x68a32de6eb7b7808ULL, 0xe9b4fe527ac8558dULL, 0xf53c6b79f162a8cbULL, 0x7cae99135f554f17ULL, 0x2c387a41f6ea09a5ULL, 0x629cbbb17a706ea3ULL, 0x99871c62d5032924ULL, 0xdbda0f3a9c6c30abULL, 0x5e92dd42e137b077ULL, 0x4ebd44f5dfd65eb9ULL, 0x00c12e131ea2b18cULL, 0xfd19e5fd0c9df246ULL, 0xf4ac8d4519009025ULL}},
};

/* Known Wieferich primes of the bases 2..30 that are not perfect powers (the OEIS values
 * the reference's selftest checks, A001220/wieferich.c:699-723). */
static const struct { u32 b; int n; u64 p[8]; } KNOWN_SOL[] = {
    {2, 2, {1093ULL, 3511ULL}},
    {3, 2, {11ULL, 1006003ULL}},
    {5, 7, {2ULL, 20771ULL, 40487ULL, 53471161ULL, 1645333507ULL, 6692367337ULL, 188748146801ULL}},
    {6, 3, {66161ULL, 534851ULL, 3152573ULL}},
    {7, 2, {5ULL, 491531ULL}},
    {10, 3, {3ULL, 487ULL, 56598313ULL}},
    {11, 1, {71ULL}},
    {12, 2, {2693ULL, 123653ULL}},
    {13, 3, {2ULL, 863ULL, 1747591ULL}},
    {14, 3, {29ULL, 353ULL, 7596952219ULL}},
    {15, 2, {29131ULL, 119327070011ULL}},
    {17, 5, {2ULL, 3ULL, 46021ULL, 48947ULL, 478225523351ULL}},
    {18, 6, {5ULL, 7ULL, 37ULL, 331ULL, 33923ULL, 1284043ULL}},
    {19, 6, {3ULL, 7ULL, 13ULL, 43ULL, 137ULL, 63061489ULL}},
    {20, 4, {281ULL, 46457ULL, 9377747ULL, 122959073ULL}},
    {21, 1, {2ULL}},
    {22, 5, {13ULL, 673ULL, 1595813ULL, 492366587ULL, 9809862296159ULL}},
    {23, 5, {13ULL, 2481757ULL, 13703077ULL, 15546404183ULL, 2549536629329ULL}},
    {24, 2, {5ULL, 25633ULL}},
    {26, 5, {3ULL, 5ULL, 71ULL, 486999673ULL, 6695256707ULL}},
    {28, 3, {3ULL, 19ULL, 23ULL}},
    {29, 1, {2ULL}},
    {30, 4, {7ULL, 160541ULL, 94727075783ULL, 303632117562967ULL}},
};
#define NKNOWN_SOL ((int)(sizeof KNOWN_SOL / sizeof *KNOWN_SOL))

static bool known_solution(u32 b, u64 p)
{
    for (int i = 0; i < NKNOWN_SOL; i++)
        if (KNOWN_SOL[i].b == b)
            for (int k = 0; k < KNOWN_SOL[i].n; k++) if (KNOWN_SOL[i].p[k] == p) return true;
    return false;
}

static void use_bases(const char *list)
{
    parse_bases(list);
    if (g_eng) {
        int rc = wf_set_bases(g_eng, g_b, g_nb);
        if (rc) die("wf_set_bases: %s", wf_strerror(rc));
    }
}

/* selftest, the reference's three parts (wieferich.c:1821-1900): GPU arithmetic against
 * shift-and-add, the "known" bases over [0, LIMIT) against the known solutions (and
 * pi(10^9)), windows around the known solutions above LIMIT, and the checksum windows
 * for the "oeis" bases (plus base 2). */
static int cmd_selftest(int argc, char **argv)
{
    u64 limit = 1000000000ULL;
    for (int i = 0; i < argc; i++) {
        if ((!strcmp(argv[i], "-t") || !strcmp(argv[i], "-T")) && i + 1 < argc) (void)parse_int(argv[++i], 1, 1024, "threads");
        else if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-') limit = parse_num(argv[i]);
        else usage();
    }
    use_bases("2");
    engine_open();
    int bad = 0;
    /* 0. GPU quotients of random odd numbers of every size against shift-and-add */
    {
        enum { NQ = 8000 };
        u64 *p = malloc(NQ * sizeof *p), *q = malloc(NQ * sizeof *q), *rh = malloc(NQ * sizeof *rh),
            *rl = malloc(NQ * sizeof *rl);
        const u64 edge[] = {3, 5, 7, 1093, 3511, 4294967291ULL, 4294967311ULL, (1ULL << 50) - 27, (1ULL << 63) - 25,
                            (1ULL << 63) + 29, 18446744073709551557ULL, UINT64_MAX};
        const u32 tb[] = {2, 3, 30, 1625, 4294967291U};
        int n = 0;
        for (size_t i = 0; i < sizeof edge / sizeof *edge; i++) p[n++] = edge[i];
        while (n < NQ) {
            int bits = 2 + (int)(rnd() % 63);
            u64 x = (rnd() >> (64 - bits)) | (1ULL << (bits - 1)) | 1;
            if (x >= 3) p[n++] = x;
        }
        int qbad = 0;
        for (size_t k = 0; k < sizeof tb / sizeof *tb; k++) {
            int rc = wf_quotients_base(g_eng, tb[k], p, q, rh, rl, NQ);
            if (rc) die("wf_quotients_base: %s", wf_strerror(rc));
            for (int i = 0; i < NQ && qbad < 5; i++) {
                const u128 rr = powmod_ref(tb[k], p[i] - 1, (u128)p[i] * p[i]);
                u64 pinv = (3 * p[i]) ^ 2;                  /* p^-1 mod 2^64 by Newton */
                for (int r = 0; r < 4; r++) pinv *= 2 - p[i] * pinv;
                const u64 qr = ((u64)rr - 1) * pinv;
                if ((((u128)rh[i] << 64) | rl[i]) != rr || q[i] != qr) {
                    qbad++;
                    fprintf(stderr, "  quotient mismatch base %u p=%" PRIu64 "\n", tb[k], p[i]);
                }
            }
        }
        fprintf(stderr, "arithmetic: %d GPU residues b^(p-1) mod p^2 for %zu bases vs shift-and-add: %s\n", NQ,
                sizeof tb / sizeof *tb, qbad ? "FAIL" : "ok");
        bad += qbad;
        free(p); free(q); free(rh); free(rl);
    }
    g_print = false;
    /* 1. the "known" bases over [0, LIMIT) against the known solutions */
    use_bases("known");
    double t = now();
    u64 chunk = limit / 16;
    if (chunk < 1000000) chunk = 1000000;
    run_scan(0, limit, chunk, NULL, 60, true);
    const bool pi_ok = limit != 1000000000ULL || S.primes == 50847534ULL;
    int found = 0, extra = 0, missing = 0;
    for (int i = 0; i < S.nhits; i++) {
        if (known_solution(S.hits[i].b, S.hits[i].p)) found++;
        else { extra++; fprintf(stderr, "  unexpected solution base %u p %" PRIu64 "\n", S.hits[i].b, S.hits[i].p); }
    }
    for (int i = 0; i < NKNOWN_SOL; i++)
        for (int k = 0; k < KNOWN_SOL[i].n; k++) {
            if (KNOWN_SOL[i].p[k] >= (u64)S.end) continue;
            bool f = false;
            for (int h = 0; h < S.nhits; h++) if (S.hits[h].b == KNOWN_SOL[i].b && S.hits[h].p == KNOWN_SOL[i].p[k]) f = true;
            if (!f) { missing++; fprintf(stderr, "  missing base %u p %" PRIu64 "\n", KNOWN_SOL[i].b, KNOWN_SOL[i].p[k]); }
        }
    fprintf(stderr, "scan [0, %" PRIu64 "), %d bases: %" PRIu64 " primes%s, %d known solutions found, %d missing, %d unexpected, %" PRIu64
                    " FLT errors (%.1f s)\n", (u64)S.end, g_nb, S.primes,
            limit == 1000000000ULL ? (pi_ok ? " (= pi(10^9))" : " (WRONG, pi(10^9) = 50847534)") : "", found, missing,
            extra, S.flt, now() - t);
    bad += missing + extra + (int)S.flt + !pi_ok;
    /* 2. windows around the known solutions above the scan */
    t = now();
    const u64 scanned = (u64)S.end;
    int wins = 0, wbad = 0;
    for (int i = 0; i < NKNOWN_SOL; i++)
        for (int k = 0; k < KNOWN_SOL[i].n; k++) {
            const u64 p = KNOWN_SOL[i].p[k];
            if (p < scanned) continue;
            char one[16];
            snprintf(one, sizeof one, "%u", KNOWN_SOL[i].b);
            use_bases(one);
            run_scan(p > 1000000 ? p - 1000000 : 0, p + 1000000, 200000, NULL, 60, true);
            const bool ok = S.nhits == 1 && S.hits[0].p == p && S.flt == 0;
            if (!ok) { wbad++; fprintf(stderr, "  window around base %u p %" PRIu64 ": %d hits\n", KNOWN_SOL[i].b, p, S.nhits); }
            wins++;
        }
    fprintf(stderr, "windows: %d known solutions above %" PRIu64 " re-found through the scan: %s (%.1f s)\n", wins,
            scanned, wbad ? "FAIL" : "ok", now() - t);
    bad += wbad;
    /* 3. checksum windows: the "oeis" bases against wieferich_cks.h, base 2 against Python integers */
    t = now();
    int cbad = 0;
    use_bases("oeis");
    for (size_t w = 0; w < sizeof CKWIN_OEIS / sizeof *CKWIN_OEIS; w++) {
        run_scan(CKWIN_OEIS[w].lo, CKWIN_OEIS[w].hi, (CKWIN_OEIS[w].hi - CKWIN_OEIS[w].lo) / 16, NULL, 60, true);
        bool ok = S.primes == CKWIN_OEIS[w].primes && S.flt == 0;
        for (int j = 0; j < g_nb; j++) if (S.cks[j] != CKWIN_OEIS[w].cks[j]) ok = false;
        if (!ok) {
            cbad++;
            fprintf(stderr, "  checksum window [%" PRIu64 ", %" PRIu64 "): primes %" PRIu64 "\n", CKWIN_OEIS[w].lo,
                    CKWIN_OEIS[w].hi, S.primes);
            for (int j = 0; j < g_nb; j++)
                if (S.cks[j] != CKWIN_OEIS[w].cks[j])
                    fprintf(stderr, "    base %u: %016" PRIx64 " expected %016" PRIx64 "\n", g_b[j], S.cks[j], CKWIN_OEIS[w].cks[j]);
        }
    }
    use_bases("2");
    for (size_t w = 0; w < sizeof CKWIN2 / sizeof *CKWIN2; w++) {
        run_scan(CKWIN2[w].lo, CKWIN2[w].hi, (CKWIN2[w].hi - CKWIN2[w].lo) / 16, NULL, 60, true);
        const bool ok = S.primes == CKWIN2[w].primes && S.flt == 0 && S.cks[0] == CKWIN2[w].cks;
        if (!ok) {
            cbad++;
            fprintf(stderr, "  checksum window [%" PRIu64 ", %" PRIu64 "), base 2: primes %" PRIu64 " cks %016" PRIx64 "\n",
                    CKWIN2[w].lo, CKWIN2[w].hi, S.primes, S.cks[0]);
        }
    }
    fprintf(stderr, "checksums: %zu windows x 17 'oeis' bases (wieferich_cks.h) and base 2 (Python integers): %s (%.1f s)\n",
            sizeof CKWIN_OEIS / sizeof *CKWIN_OEIS, cbad ? "FAIL" : "ok", now() - t);
    bad += cbad;
    fprintf(stderr, "selftest %s\n", bad ? "FAILED" : "passed");
    wf_close(g_eng);
    return bad ? 1 : 0;
}

/* num X [X ...]: the GPU's base-2 strong test, strong Lucas test and BPSW verdict of each X
 * (the reference's `a350826_cuda num`, A350826/cuda/a350826_cuda.cu:2458-2483). */
static int cmd_num(int argc, char **argv)
{
    u64 xs[256];
    int n = 0;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-' && n < 256) xs[n++] = parse_num(argv[i]);
        else usage();
    }
    if (!n) usage();
    g_b[0] = 2;
    g_nb = 1;
    engine_open();
    uint8_t f[256];
    const int rc = wf_bpsw(g_eng, xs, f, (u64)n);
    if (rc) die("wf_bpsw: %s", wf_strerror(rc));
    for (int i = 0; i < n; i++)
        printf("%" PRIu64 ": gpu sprp2 %d slprp %d prime %d\n", xs[i], f[i] & 1, (f[i] >> 1) & 1, (f[i] >> 2) & 1);
    wf_close(g_eng);
    return 0;
}

/* dcheck CHUNKLOG BASES [STEP] [THREADS] [OUT] [PID]: A001220/dcheck.sh on the GPU.
 * Every STEP-th line (default 16) of a chunk log is recomputed with an engine set up
 * differently from a scan's: prove mode for every piece (small-prime sieve + BPSW instead
 * of the bucket sieve), 16 KiB segments instead of 64 KiB, and for the single base 2 the
 * 32-bit limb Montgomery ladders instead of the FP64 p-adic / radix-2^28 arithmetic.  The
 * recomputed line must equal the logged one byte for byte; "MATCH lo hi TIME" or
 * "MISMATCH lo hi gpu: LINE check: LINE" is appended to OUT (default CHUNKLOG.dcheck).
 * Chunks already in OUT are skipped, so a run can be resumed; with PID, new chunks are
 * awaited while that process lives (dcheck.sh:12-45).  THREADS is accepted and ignored.
 * Exit status 0 when every chunk checked in this run matched, 1 otherwise. */
static bool out_has(const char *out, u64 lo, const char *hi)
{
    FILE *f = fopen(out, "r");
    if (!f) return false;
    char key[64], line[8192];
    snprintf(key, sizeof key, " %" PRIu64 " %s ", lo, hi);
    bool found = false;
    while (!found && fgets(line, sizeof line, f)) found = strstr(line, key) != NULL;
    fclose(f);
    return found;
}

static int cmd_dcheck(int argc, char **argv)
{
    const char *pos[6] = {0};
    int npos = 0;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-' && npos < 6) pos[npos++] = argv[i];
        else usage();
    }
    if (npos < 2) usage();
    const char *logp = pos[0];
    parse_bases(pos[1]);
    const u64 step = npos > 2 ? parse_num(pos[2]) : 16;
    if (step == 0) die("STEP must be positive");
    if (npos > 3) (void)parse_int(pos[3], 1, 1024, "threads");
    char outbuf[4096];
    snprintf(outbuf, sizeof outbuf, "%s.dcheck", logp);
    const char *outp = npos > 4 ? pos[4] : outbuf;
    const long pid = npos > 5 ? (long)parse_num(pos[5]) : 0;
    /* the recomputation takes paths a scan does not: prove mode, small segments, one ladder
     * per base (a scan derives composite bases from their prime factors), and for base 2
     * alone the 32-bit limb ladders */
    g_cfg.prove = WF_PROVE_ON;
    g_cfg.seg_bytes = 16384;
    g_cfg.derive = WF_DERIVE_OFF;
    if (g_nb == 1 && g_b[0] == 2) g_cfg.arith = 1;
    engine_open();
    FILE *lf = fopen(logp, "r");
    if (!lf) die("cannot open %s: %s", logp, strerror(errno));
    const u64 EVCAP = 1 << 16;
    wf_event *ev = malloc(EVCAP * sizeof *ev);
    if (!ev) die("out of memory");
    u64 k = 0, checked = 0, bad = 0;
    char line[64 + WF_MAXB * 32], mine[64 + WF_MAXB * 32];
    for (;;) {
        bool progressed = false;
        long where = ftell(lf);
        while (fgets(line, sizeof line, lf)) {
            size_t n = strlen(line);
            if (n == 0 || line[n - 1] != '\n') {           /* a line still being written */
                fseek(lf, where, SEEK_SET);
                break;
            }
            line[--n] = 0;
            where = ftell(lf);
            k++;
            if (k % step) continue;
            char los[32], his[32];
            if (sscanf(line, "C %31s %31s", los, his) != 2) die("%s: not a chunk-log line: %s", logp, line);
            const u64 lo = parse_num(los);
            const u128 hi = parse_num128(his);
            if (hi <= lo || hi > LIMIT128) die("%s: bad range in: %s", logp, line);
            if (out_has(outp, lo, his)) continue;
            wf_result r;
            int rc = wf_scan(g_eng, lo, hi >= LIMIT128 ? 0 : (u64)hi, 0, &r, ev, EVCAP, 0, NULL, 0);
            if (rc) die("chunk [%" PRIu64 ", %s): %s", lo, his, wf_strerror(rc));
            if (r.flt) snprintf(mine, sizeof mine, "BAD range [%" PRIu64 ", %s) flt %" PRIu64, lo, his, r.flt);
            else format_chunk_line(mine, sizeof mine, lo, hi, r.primes, r.cks);
            FILE *of = fopen(outp, "a");
            if (!of) die("cannot open %s: %s", outp, strerror(errno));
            if (!strcmp(mine, line)) {
                char ts[32];
                const time_t t = time(NULL);
                strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%S", localtime(&t));
                fprintf(of, "MATCH %" PRIu64 " %s %s\n", lo, his, ts);
            } else {
                fprintf(of, "MISMATCH %" PRIu64 " %s gpu: %s check: %s\n", lo, his, line, mine);
                bad++;
            }
            fclose(of);
            checked++;
            progressed = true;
        }
        clearerr(lf);
        if (pid > 0 && kill((pid_t)pid, 0) == 0) {
            if (!progressed) sleep(60);
        } else if (!progressed) {
            break;
        }
    }
    fclose(lf);
    free(ev);
    wf_close(g_eng);
    fprintf(stderr, "dcheck: %" PRIu64 " chunks checked, %" PRIu64 " mismatches\n", checked, bad);
    return bad ? 1 : 0;
}

/* stats [LO] HI: the sieve's statistics for the repository's other prime consumers: the
 * twin pairs of A007508 (p, p + 2 both prime, p < HI), the largest gap, and the run lengths
 * of A158939 (L(p) = number of strictly increasing gaps after p, A158939/a158939.c:19-24)
 * with the first prime of each. */
static int cmd_stats(int argc, char **argv)
{
    u128 pos[2];
    int npos = 0;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "-g") && i + 1 < argc) g_dev = parse_int(argv[++i], 0, 64, "device");
        else if (argv[i][0] != '-' && npos < 2) pos[npos++] = parse_num128(argv[i]);
        else usage();
    }
    if (npos == 0) usage();
    const u128 lo = npos == 2 ? pos[0] : 0, hi = npos == 2 ? pos[1] : pos[0];
    if (hi <= lo) die("END must be larger than START");
    if (hi > (u128)WF_STATS_MAXHI) die("END must be <= 2^64 - 2^20 (the gaps after the last primes are needed)");
    g_b[0] = 2;
    g_nb = 1;
    engine_open();
    wf_stats st;
    const double t0 = now();
    const int rc = wf_sieve_stats(g_eng, (u64)lo, (u64)hi, &st);
    if (rc) die("wf_sieve_stats: %s", wf_strerror(rc));
    printf("STATS [%" PRIu64 ", %" PRIu64 ") primes %" PRIu64 " twins %" PRIu64 " max_gap %" PRIu64 " after %" PRIu64 "\n",
           (u64)lo, (u64)hi, st.primes, st.twins, st.max_gap, st.max_gap_p);
    for (int n = 1; n < 64; n++)
        if (st.runs[n]) printf("run %2d count %" PRIu64 " first %" PRIu64 "\n", n, st.runs[n], st.first[n]);
    fprintf(stderr, "%.3f s\n", now() - t0);
    wf_close(g_eng);
    return 0;
}

int main(int argc, char **argv)
{
    stderr_tty = isatty(2);
    if (argc < 2) usage();
    const char *cmd = argv[1];
    if (!strcmp(cmd, "scan")) return cmd_scan(argc - 2, argv + 2);
    if (!strcmp(cmd, "check")) return cmd_check(argc - 2, argv + 2);
    if (!strcmp(cmd, "bench")) return cmd_bench(argc - 2, argv + 2);
    if (!strcmp(cmd, "selftest")) return cmd_selftest(argc - 2, argv + 2);
    if (!strcmp(cmd, "num")) return cmd_num(argc - 2, argv + 2);
    if (!strcmp(cmd, "dcheck")) return cmd_dcheck(argc - 2, argv + 2);
    if (!strcmp(cmd, "stats")) return cmd_stats(argc - 2, argv + 2);
    usage();
    return 2;
}
