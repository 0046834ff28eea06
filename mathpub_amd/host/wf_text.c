/*
 * wf_text.c -- number and base-list grammar, chunk log, checkpoint and SUMMARY text of the
 * wf host program (declarations and format references in wf_text.h).
 */
#define _GNU_SOURCE
#include "wf_text.h"

#include <inttypes.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef wt_u128 u128;

/* ------------------------------------------------------------------ numbers */

static bool is_digit(char c) { return c >= '0' && c <= '9'; }

/* v = v * m + a, kept at or below 2^64. */
static bool grow_checked(u128 *v, unsigned m, unsigned a)
{
    if (*v > (WT_TWO64 - a) / m) return false;
    *v = *v * m + a;
    return true;
}

/* s[0..n) as a decimal integer <= 2^64; false for an empty run, a non-digit or a larger value. */
static bool decimal(const char *s, size_t n, u128 *out)
{
    u128 v = 0;
    if (n == 0) return false;
    for (size_t i = 0; i < n; i++)
        if (!is_digit(s[i]) || !grow_checked(&v, 10, (unsigned)(s[i] - '0'))) return false;
    *out = v;
    return true;
}

static const char *find_char(const char *s, size_t n, const char *set)
{
    for (size_t i = 0; i < n; i++)
        if (strchr(set, s[i])) return s + i;
    return NULL;
}

/* B^E by repeated squaring, every intermediate kept at or below 2^64. */
static bool power(u128 b, u128 e, u128 *out)
{
    if (b <= 1) {                 /* 0^0 = 1^E = 1, 0^E = 0 */
        *out = (e == 0 || b == 1) ? 1 : 0;
        return true;
    }
    if (e > 64) return false;     /* 2^65 and anything larger */
    u128 r = 1;
    for (u128 k = 0; k < e; k++) {
        if (r > WT_TWO64 / b) return false;
        r *= b;
    }
    *out = r;
    return true;
}

/* M e E: M = digits with at most one decimal point and at least one digit, E = digits.
 * The value M * 10^E must be an integer <= 2^64. */
static bool scientific(const char *s, size_t n, size_t epos, u128 *out)
{
    const char *m = s, *x = s + epos + 1;
    const size_t mlen = epos, xlen = n - epos - 1;
    u128 ex;
    if (!decimal(x, xlen, &ex)) return false;
    /* digits of M with the point removed; `after` counts those after the point */
    size_t ndig = 0, after = 0;
    bool point = false;
    for (size_t i = 0; i < mlen; i++) {
        if (m[i] == '.') {
            if (point) return false;
            point = true;
        } else if (is_digit(m[i])) {
            ndig++;
            if (point) after++;
        } else {
            return false;
        }
    }
    if (ndig == 0) return false;
    /* zeros at the end of the fraction do not change the value */
    size_t drop = 0;
    if (point)
        for (size_t i = mlen; i-- > 0 && m[i] != '.' && m[i] == '0';) drop++;
    u128 v = 0;
    size_t used = 0;
    for (size_t i = 0; i < mlen && used < ndig - drop; i++) {
        if (m[i] == '.') continue;
        if (!grow_checked(&v, 10, (unsigned)(m[i] - '0'))) return false;
        used++;
    }
    after -= drop;
    if (v == 0) {
        *out = 0;
        return true;
    }
    /* value = v * 10^(ex - after) */
    if (ex >= after) {
        for (u128 k = after; k < ex; k++)
            if (!grow_checked(&v, 10, 0)) return false;
    } else {
        for (u128 k = ex; k < after; k++) {
            if (v % 10) return false;
            v /= 10;
        }
    }
    *out = v;
    return true;
}

/* One term (no + or - operator). */
static bool term(const char *s, size_t n, u128 *out)
{
    if (n == 0 || n > WT_TERM_MAX) return false;
    const char *c = find_char(s, n, "^");
    if (c) {
        u128 b, e;
        const size_t k = (size_t)(c - s);
        return decimal(s, k, &b) && decimal(c + 1, n - k - 1, &e) && power(b, e, out);
    }
    if ((c = find_char(s, n, "eE"))) return scientific(s, n, (size_t)(c - s), out);
    return decimal(s, n, out);
}

int wt_number(const char *s, u128 *out)
{
    const size_t n = strlen(s);
    /* the operator: the last + or - after the first character that does not follow an
     * exponent's e (so "1e-5" is one term, and rejected) */
    size_t op = 0;
    for (size_t i = 1; i < n; i++)
        if ((s[i] == '+' || s[i] == '-') && s[i - 1] != 'e' && s[i - 1] != 'E') op = i;
    if (!op) return term(s, n, out) ? WT_OK : WT_BAD;
    u128 a, b;
    if (!term(s, op, &a) || !term(s + op + 1, n - op - 1, &b)) return WT_BAD;
    if (s[op] == '+') {
        if (a + b > WT_TWO64) return WT_OVERFLOW;
        *out = a + b;
    } else {
        if (b > a) return WT_NEGATIVE;
        *out = a - b;
    }
    return WT_OK;
}

/* ------------------------------------------------------------------ bases */

const uint32_t WT_OEIS[17] = {3, 5, 6, 7, 10, 12, 13, 14, 15, 17, 18, 19, 20, 22, 23, 26, 30};
const uint32_t WT_KNOWN[23] = {2, 3, 5, 6, 7, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 23, 24, 26, 28, 29, 30};

const char *wt_anum(uint32_t b)
{
    static const struct { u32 b; const char *a; } T[] = {
        {2, "A001220"}, {3, "A014127"}, {5, "A123692"}, {6, "A212583"}, {7, "A123693"}, {10, "A045616"},
        {12, "A111027"}, {13, "A128667"}, {14, "A234810"}, {15, "A242741"}, {17, "A128668"}, {18, "A244260"},
        {19, "A090968"}, {20, "A242982"}, {22, "A298951"}, {23, "A128669"}, {26, "A306255"}, {30, "A306256"}};
    for (size_t i = 0; i < sizeof T / sizeof *T; i++)
        if (T[i].b == b) return T[i].a;
    return NULL;
}

/* Add b to the sorted set out[0..*nb); false if the set is full. */
static bool add_base(uint32_t *out, int *nb, u32 b)
{
    int i = 0;
    while (i < *nb && out[i] < b) i++;
    if (i < *nb && out[i] == b) return true;
    if (*nb >= WT_MAXB) return false;
    memmove(out + i + 1, out + i, (size_t)(*nb - i) * sizeof *out);
    out[i] = b;
    (*nb)++;
    return true;
}

static bool add_list(uint32_t *out, int *nb, const u32 *list, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!add_base(out, nb, list[i])) return false;
    return true;
}

/* A number of a base list as a u64 value; messages as the command line reports them. */
static bool base_number(const char *s, u64 *v, char *err, size_t errlen)
{
    u128 x;
    const int rc = wt_number(s, &x);
    if (rc == WT_OK && x < WT_TWO64) {
        *v = (u64)x;
        return true;
    }
    snprintf(err, errlen, "%s '%s'", rc == WT_NEGATIVE ? "negative result in" : rc == WT_BAD ? "bad number" : "overflow in",
             s);
    return false;
}

bool wt_bases(const char *s, uint32_t *out, int *nb, char *err, size_t errlen)
{
    *nb = 0;
    if (strlen(s) > WT_BASES_MAX) {
        snprintf(err, errlen, "base list too long");
        return false;
    }
    const char *p = s;
    while (*p) {
        const char *q = strchr(p, ',');
        const size_t len = q ? (size_t)(q - p) : strlen(p);
        char item[WT_BASES_MAX + 1];
        memcpy(item, p, len);
        item[len] = 0;
        p += len + (q ? 1 : 0);
        if (len == 0) continue;              /* empty items are skipped */
        bool ok = true;
        if (!strcmp(item, "oeis")) {
            ok = add_list(out, nb, WT_OEIS, 17);
        } else if (!strcmp(item, "known")) {
            ok = add_list(out, nb, WT_KNOWN, 23);
        } else {
            /* a range LO-HI splits at its first '-' */
            char *dash = strchr(item, '-');
            u64 lo, hi;
            if (dash) {
                *dash = 0;
                if (!base_number(item, &lo, err, errlen) || !base_number(dash + 1, &hi, err, errlen)) return false;
                *dash = '-';
            } else {
                if (!base_number(item, &lo, err, errlen)) return false;
                hi = lo;
            }
            if (lo < 2 || hi < lo || hi > UINT32_MAX) {
                snprintf(err, errlen, "bad base or range '%s'", item);
                return false;
            }
            for (u64 b = lo; b <= hi && ok; b++) ok = add_base(out, nb, (u32)b);
        }
        if (!ok) {
            snprintf(err, errlen, "too many bases (max %d)", WT_MAXB);
            return false;
        }
    }
    if (*nb == 0) {
        snprintf(err, errlen, "between 1 and %d bases are supported", WT_MAXB);
        return false;
    }
    return true;
}

void wt_join_bases(char *buf, size_t n, const uint32_t *b, int nb)
{
    size_t k = 0;
    if (n) buf[0] = 0;
    for (int j = 0; j < nb && k + 1 < n; j++) {
        const int w = snprintf(buf + k, n - k, j ? ",%u" : "%u", b[j]);
        if (w < 0) break;
        k += (size_t)w;
    }
}

/* ------------------------------------------------------------------ lines */

void wt_u128_str(char *buf, size_t n, u128 v)
{
    if (v >> 64) snprintf(buf, n, "18446744073709551616");   /* only 2^64 occurs */
    else snprintf(buf, n, "%" PRIu64, (u64)v);
}

void wt_chunk_line(char *buf, size_t n, uint64_t lo, u128 hi, uint64_t primes, const uint32_t *b,
                   const uint64_t *cks, int nb)
{
    char his[24];
    wt_u128_str(his, sizeof his, hi);
    size_t k = (size_t)snprintf(buf, n, "C %" PRIu64 " %s %" PRIu64, lo, his, primes);
    for (int j = 0; j < nb && k < n; j++) k += (size_t)snprintf(buf + k, n - k, " %u:%016" PRIx64, b[j], cks[j]);
}

void wt_duration(char *buf, size_t n, double s)
{
    if (!(s >= 0) || s > 1e12) {
        snprintf(buf, n, "?");
        return;
    }
    const long t = (long)s, d = t / 86400, h = t / 3600 % 24, m = t / 60 % 60, sec = t % 60;
    if (d) snprintf(buf, n, "%ldd%02ldh%02ldm", d, h, m);
    else if (h) snprintf(buf, n, "%ldh%02ldm%02lds", h, m, sec);
    else if (m) snprintf(buf, n, "%ldm%02lds", m, sec);
    else snprintf(buf, n, "%.1fs", s);
}

/* ------------------------------------------------------------------ checkpoint */

#define CK_MAGIC "wieferich checkpoint 1"

/* A u64 field value: decimal digits only, at most 2^64 - 1. */
static bool field_u64(const char *s, u64 *v)
{
    u128 x;
    if (!decimal(s, strlen(s), &x) || x >= WT_TWO64) return false;
    *v = (u64)x;
    return true;
}

static bool field_hex(const char *s, u64 *v)
{
    const size_t n = strlen(s);
    if (n == 0 || n > 16) return false;
    u64 x = 0;
    for (size_t i = 0; i < n; i++) {
        const char c = s[i];
        const int d = is_digit(c) ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10 : (c >= 'A' && c <= 'F') ? c - 'A' + 10 : -1;
        if (d < 0) return false;
        x = x << 4 | (u64)d;
    }
    *v = x;
    return true;
}

/* Split a line into at most `max` space-separated words (in place). */
static int words(char *line, char **w, int max)
{
    int n = 0;
    for (char *p = line; *p;) {
        while (*p == ' ' || *p == '\t') *p++ = 0;
        if (!*p) break;
        if (n == max) return max + 1;
        w[n++] = p;
        while (*p && *p != ' ' && *p != '\t') p++;
    }
    return n;
}

void wt_ckpt_free(wt_ckpt *c)
{
    free(c->hits);
    c->hits = NULL;
    c->nhits = 0;
}

int wt_ckpt_read(FILE *f, const char *path, wt_ckpt *c, char *err, size_t errlen)
{
    memset(c, 0, sizeof *c);
    char *line = NULL;
    size_t cap = 0, hcap = 0;
    ssize_t len;
    int rc = WT_CK_TRUNC;
    bool first = true;
    while ((len = getline(&line, &cap, f)) >= 0) {
        while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
        if (first) {
            first = false;
            if (strncmp(line, CK_MAGIC, strlen(CK_MAGIC)) != 0) {
                rc = WT_CK_NOTCK;
                break;
            }
            continue;
        }
        char *w[5];
        const int nw = words(line, w, 4);
        if (nw == 0) continue;
        const char *key = w[0];
        bool ok = true;
        if (!strcmp(key, "end") && nw == 1) {        /* the terminator */
            rc = 0;
            break;
        }
        if (!strcmp(key, "bases")) {
            ok = nw == 2 && strlen(w[1]) <= WT_BASES_MAX;
            if (ok) strcpy(c->bases, w[1]);
        } else if (!strcmp(key, "end")) {
            ok = nw == 2 && decimal(w[1], strlen(w[1]), &c->end);
        } else if (!strcmp(key, "cks")) {
            u64 b;
            ok = nw == 3 && field_u64(w[1], &b) && b <= UINT32_MAX && field_hex(w[2], &c->cks[c->ncks]);
            if (ok && c->ncks == WT_MAXB) ok = false;
            if (ok) c->cks_base[c->ncks++] = (u32)b;
        } else if (!strcmp(key, "hit")) {
            u64 b;
            wt_hit h;
            ok = nw == 4 && field_u64(w[1], &b) && b <= UINT32_MAX && field_u64(w[2], &h.p) && field_u64(w[3], &h.chunk);
            if (ok) {
                if (c->nhits == hcap) {
                    hcap = hcap ? 2 * hcap : 16;
                    wt_hit *nh = realloc(c->hits, hcap * sizeof *nh);
                    if (!nh) {
                        rc = WT_CK_NOMEM;
                        snprintf(err, errlen, "out of memory reading %s", path);
                        break;
                    }
                    c->hits = nh;
                }
                h.b = (u32)b;
                c->hits[c->nhits++] = h;
            }
        } else {
            static const char *const names[] = {"start", "chunk", "frontier", "primes", "tests", "near", "flt"};
            u64 *const dst[] = {&c->start, &c->chunk, &c->frontier, &c->primes, &c->tests, &c->near, &c->flt};
            for (size_t i = 0; i < sizeof names / sizeof *names; i++)
                if (!strcmp(key, names[i])) ok = nw == 2 && field_u64(w[1], dst[i]);
            /* any other line is ignored (a later version may add fields) */
        }
        if (!ok) {
            rc = WT_CK_LINE;
            snprintf(err, errlen, "%s: bad %s line", path, key);
            break;
        }
    }
    free(line);
    if (first && rc == WT_CK_TRUNC) rc = WT_CK_NOTCK;    /* empty file */
    if (rc == WT_CK_NOTCK) snprintf(err, errlen, "%s is not a wieferich checkpoint", path);
    else if (rc == WT_CK_TRUNC) snprintf(err, errlen, "%s is truncated", path);
    if (rc) wt_ckpt_free(c);
    return rc;
}

bool wt_ckpt_write(FILE *f, const wt_ckpt *c)
{
    char end[24];
    wt_u128_str(end, sizeof end, c->end);
    fprintf(f, CK_MAGIC "\nbases %s\nstart %" PRIu64 "\nchunk %" PRIu64 "\nend %s\nfrontier %" PRIu64
               "\nprimes %" PRIu64 "\ntests %" PRIu64 "\nnear %" PRIu64 "\nflt %" PRIu64 "\n",
            c->bases, c->start, c->chunk, end, c->frontier, c->primes, c->tests, c->near, c->flt);
    for (int j = 0; j < c->ncks; j++) fprintf(f, "cks %u %016" PRIx64 "\n", c->cks_base[j], c->cks[j]);
    for (size_t i = 0; i < c->nhits; i++)
        if (c->hits[i].chunk < c->frontier)
            fprintf(f, "hit %u %" PRIu64 " %" PRIu64 "\n", c->hits[i].b, c->hits[i].p, c->hits[i].chunk);
    fputs("end\n", f);
    return !ferror(f);
}

/* ------------------------------------------------------------------ summary */

void wt_summary(FILE *f, uint64_t start, u128 pos, uint64_t primes, uint64_t tests, uint64_t flt, long long near_T,
                uint64_t near, const uint32_t *b, int nb, const uint64_t *cks, const uint64_t *sumq, const wt_hit *hits,
                size_t nhits, uint64_t frontier)
{
    char ps[24];
    wt_u128_str(ps, sizeof ps, pos);
    fprintf(f, "SUMMARY [%" PRIu64 ", %s) primes %" PRIu64 " tests %" PRIu64 " flt_errors %" PRIu64, start, ps, primes,
            tests, flt);
    if (near_T > 0) fprintf(f, " near(|A|<=%lld) %" PRIu64, near_T, near);
    fputc('\n', f);
    for (int j = 0; j < nb; j++) {
        const char *a = wt_anum(b[j]);
        fprintf(f, "base %3u %-8s cks %016" PRIx64 "  solutions:", b[j], a ? a : "", cks[j]);
        size_t shown = 0;
        for (size_t i = 0; i < nhits; i++)
            if (hits[i].b == b[j] && hits[i].chunk < frontier) {
                fprintf(f, " %" PRIu64, hits[i].p);
                shown++;
            }
        fputs(shown ? "\n" : " none\n", f);
        if (sumq) fprintf(f, "sumq base %u %016" PRIx64 "\n", b[j], sumq[j]);
    }
    fflush(f);
}
