/*
 * wf_oracle.c -- CPU oracle for the Wieferich sieve -> Fermat-quotient path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP engine and the CPU baseline
 * timed by bench.py.  The product never links or calls this file.  Every function
 * cites the part of A001220/wieferich.c (jeffsponaugle/Mathpub) whose algorithm it
 * restates; see wf_oracle.h for the overview.
 *
 * Build: make -C oracle   (gcc -O2 -std=gnu11 -pthread -shared -fPIC).
 */
#define _GNU_SOURCE
#include "wf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

#define GOLD 0x9E3779B97F4A7C15ULL      /* wieferich.c:240 */
#define LANES 8                         /* wieferich.c:233 */
#define BATCH 64                        /* wieferich.c:235 */
#define MAXW 5                          /* wieferich.c:237 */
#define PMAX (1ULL << 63)               /* fast path limit, wieferich.c:241 */

/* ------------------------------------------------------------------ */
/* helpers                                                              */
/* ------------------------------------------------------------------ */

/* splitmix64 finalizer: wieferich.c:452-458, verify_wieferich.py:30-34. */
u64 wfo_mix64(u64 z)
{
    z += GOLD;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

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

/* One strong probable-prime round to base a for odd n > 2 with n - 1 = d 2^s. */
static bool sprp_round(u64 n, u64 a, u64 d, int s)
{
    u64 x = powmod64(a, d, n);
    if (x == 1 || x == n - 1) return true;
    for (int r = 1; r < s; r++) {
        x = mulmod64(x, x, n);
        if (x == n - 1) return true;
    }
    return false;
}

/* Deterministic primality below 2^64: trial division by the primes <= 37, then
 * strong tests to the same 12 bases (wieferich.c:500-522). */
int wfo_is_prime64(u64 n)
{
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++)
        if (!sprp_round(n, small[i], d, s)) return 0;
    return 1;
}

/* Base-2 strong probable-prime test (the a = 2 round of the test above). */
int wfo_sprp2(u64 n)
{
    if (n < 3 || !(n & 1)) return n == 2;
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    return sprp_round(n, 2, d, s);
}

/* ------------------------------------------------------------------ */
/* arithmetic mod N = p^2, p < 2^63 (wieferich.c:524-645)               */
/* ------------------------------------------------------------------ */

typedef struct {
    u128 N, one;
    double invN;
    u64 p, pinv, ninv;
    bool full;              /* p >= 2^63: N may reach 2^128, see the full path below */
} pctx_t;

/* Per-prime constants: pinv by Newton from 3p^2 (5 -> 80 bits), ninv = -pinv^2,
 * one = 2^128 mod N, invN (wieferich.c:546-564). */
static void pctx_init(pctx_t *c, u64 p)
{
    u64 x = (3 * p) ^ 2;
    for (int i = 0; i < 4; i++) x *= 2 - p * x;
    c->p = p;
    c->pinv = x;
    c->N = (u128)p * p;
    c->ninv = 0 - x * x;
    c->invN = 1.0 / (double)c->N;
    c->full = p >= PMAX;
    if (!c->full && (c->N >> 80)) {
        u64 q = (u64)(0x1p128 * c->invN);
        u128 r = (u128)0 - (u128)q * c->N;
        if ((__int128)r < 0) r += c->N;
        else if (r >= c->N) r -= c->N;
        c->one = r;
    } else {
        c->one = ((u128)0 - c->N) % c->N;
    }
}

/* Montgomery square a^2 / 2^128 mod N for a < N < 2^126 (wieferich.c:573-602). */
static inline u128 msqr(u128 a, const pctx_t *m)
{
    const u64 a0 = (u64)a, a1 = (u64)(a >> 64);
    const u64 n0 = (u64)m->N, n1 = (u64)(m->N >> 64);
    const u128 p00 = (u128)a0 * a0, p01 = (u128)a0 * a1, p11 = (u128)a1 * a1;
    const u128 d = p01 << 1;
    const u64 t0 = (u64)p00;
    u128 acc = (u128)(u64)(p00 >> 64) + (u64)d;
    u64 t1 = (u64)acc;
    acc = (acc >> 64) + (u64)(d >> 64) + (u64)p11;
    u64 t2 = (u64)acc;
    u64 t3 = (u64)(acc >> 64) + (u64)(p11 >> 64);
    u64 q = t0 * m->ninv;
    acc = (u128)q * n0 + t0;
    acc = (u128)q * n1 + t1 + (u64)(acc >> 64);
    t1 = (u64)acc;
    acc = (u128)t2 + (u64)(acc >> 64);
    t2 = (u64)acc;
    t3 += (u64)(acc >> 64);
    q = t1 * m->ninv;
    acc = (u128)q * n0 + t1;
    acc = (u128)q * n1 + t2 + (u64)(acc >> 64);
    const u64 r0 = (u64)acc;
    acc = (u128)t3 + (u64)(acc >> 64);
    const u128 r = ((u128)(u64)acc << 64) | r0;
    return r >= m->N ? r - m->N : r;
}

/* x * c mod N for c < 2^32 with a double-precision quotient estimate
 * (wieferich.c:612-624). */
static inline u128 msmall(u128 x, u64 c, const pctx_t *m)
{
    const u128 lo = (u128)(u64)x * c;
    const u128 hi = (u128)(u64)(x >> 64) * c;
    const u128 t = lo + (hi << 64);
    const u64 top = (u64)(hi >> 64) + (t < lo);
    const double td = (double)top * 0x1p128 + (double)(u64)(t >> 64) * 0x1p64 + (double)(u64)t;
    const u64 q = (u64)(td * m->invN);
    u128 r = t - (u128)q * m->N;
    if ((__int128)r < 0) r += m->N;
    else if (r >= m->N) r -= m->N;
    return r;
}

/* Leave Montgomery form, x / 2^128 mod N (wieferich.c:632-645). */
static inline u128 redc(u128 x, const pctx_t *m)
{
    const u64 n0 = (u64)m->N, n1 = (u64)(m->N >> 64);
    u64 t0 = (u64)x, t1 = (u64)(x >> 64);
    u64 q = t0 * m->ninv;
    u128 acc = (u128)q * n0 + t0;
    acc = (u128)q * n1 + t1 + (u64)(acc >> 64);
    t0 = (u64)acc;
    t1 = (u64)(acc >> 64);
    q = t0 * m->ninv;
    acc = (u128)q * n0 + t0;
    acc = (u128)q * n1 + t1 + (u64)(acc >> 64);
    return acc >= m->N ? acc - m->N : acc;
}

/* ------------------------------------------------------------------ */
/* full 128-bit path, 2^63 <= p < 2^64 (beyond the reference)           */
/* N = p^2 may use all 128 bits, so REDC keeps a 129th (carry) bit and   */
/* every result is reduced to [0, N).                                    */
/* ------------------------------------------------------------------ */

/* (a b + Q N) / 2^128 mod N for a, b < N < 2^128 with R = 2^128. */
static u128 mmul_full(u128 a, u128 b, const pctx_t *m)
{
    const u64 a0 = (u64)a, a1 = (u64)(a >> 64), b0 = (u64)b, b1 = (u64)(b >> 64);
    const u64 n0 = (u64)m->N, n1 = (u64)(m->N >> 64);
    u64 t[5];
    u128 acc = (u128)a0 * b0;
    t[0] = (u64)acc;
    u128 mid1 = (u128)a0 * b1, mid2 = (u128)a1 * b0, h = (u128)a1 * b1;
    acc = (acc >> 64) + (u64)mid1 + (u64)mid2;
    t[1] = (u64)acc;
    acc = (acc >> 64) + (mid1 >> 64) + (mid2 >> 64) + (u64)h;
    t[2] = (u64)acc;
    acc = (acc >> 64) + (h >> 64);
    t[3] = (u64)acc;
    t[4] = (u64)(acc >> 64);                    /* 0: a b < 2^256 */
    for (int i = 0; i < 2; i++) {
        const u64 q = t[i] * m->ninv;
        acc = (u128)q * n0 + t[i];
        acc = (u128)q * n1 + t[i + 1] + (u64)(acc >> 64);
        t[i + 1] = (u64)acc;
        u64 c = (u64)(acc >> 64);
        for (int k = i + 2; k < 5 && c; k++) {
            acc = (u128)t[k] + c;
            t[k] = (u64)acc;
            c = (u64)(acc >> 64);
        }
    }
    u128 r = ((u128)t[3] << 64) | t[2];
    if (t[4] || r >= m->N) r -= m->N;           /* result < 2N: one subtraction */
    return r;
}

/* 2x mod N for x < N < 2^128 (the carry out of 2x is the 129th bit). */
static inline u128 mdbl_full(u128 x, const pctx_t *m)
{
    const bool c = (x >> 127) != 0;
    x <<= 1;
    if (c || x >= m->N) x -= m->N;
    return x;
}

/* b^e mod N, plain form, for the full path: left-to-right binary; multiplication by
 * b is Montgomery multiplication by bR mod N (bR built from one by doublings and
 * additions). */
static u128 powmod_full(u64 b, u64 e, const pctx_t *m)
{
    u128 bR = 0, add = m->one;                  /* bR = b * one mod N by binary */
    for (u64 k = b; k; k >>= 1) {
        if (k & 1) {
            u128 s = bR + add;
            bool c = s < bR;
            if (c || s >= m->N) s -= m->N;
            bR = s;
        }
        add = mdbl_full(add, m);
    }
    u128 x = bR;
    int top = 63 - __builtin_clzll(e);
    for (int i = top - 1; i >= 0; i--) {
        x = mmul_full(x, x, m);
        if ((e >> i) & 1) x = (b == 2) ? mdbl_full(x, m) : mmul_full(x, bR, m);
    }
    return mmul_full(x, 1, m);                  /* leave Montgomery form */
}

/* ------------------------------------------------------------------ */
/* shift-and-add reference (wieferich.c:654-683), for m < 2^128          */
/* ------------------------------------------------------------------ */

static u128 addmod_ref(u128 a, u128 b, u128 m)
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
        if (b & 1) r = addmod_ref(r, a, m);
        a = addmod_ref(a, a, m);
        b >>= 1;
    }
    return r;
}

void wfo_powmod_ref(u32 b, u64 e, u64 m_hi, u64 m_lo, u64 *r_hi, u64 *r_lo)
{
    const u128 m = ((u128)m_hi << 64) | m_lo;
    u128 r = 1 % m, x = b % m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod_ref(r, x, m);
        x = mulmod_ref(x, x, m);
    }
    *r_hi = (u64)(r >> 64);
    *r_lo = (u64)r;
}

/* ------------------------------------------------------------------ */
/* bases and the lane kernel (wieferich.c:760-949)                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int nb;
    u32 b[WFO_MAXB];
    int w[WFO_MAXB];
    u64 pw[WFO_MAXB][1 << MAXW];
    int ncls, cls_w[MAXW + 1], cls_n[MAXW + 1], cls_j[MAXW + 1][WFO_MAXB];
} bases_t;

/* Largest w <= 5 with b^(2^w - 1) < 2^32 (wieferich.c:768-780). */
static int window_for(u32 b)
{
    for (int w = MAXW; w > 1; w--) {
        u64 x = 1;
        bool ok = true;
        for (int i = 0; i < (1 << w) - 1 && ok; i++) {
            x *= b;
            if (x >> 32) ok = false;
        }
        if (ok) return w;
    }
    return 1;
}

/* Bases in the caller's order (the caller passes them sorted and distinct), power
 * tables and window classes from large w to small (wieferich.c:780-801). */
static int bases_setup(bases_t *B, const u32 *list, int n)
{
    if (n < 1 || n > WFO_MAXB) return -1;
    memset(B, 0, sizeof *B);
    for (int i = 0; i < n; i++) {
        if (list[i] < 2) return -1;
        B->b[i] = list[i];
        B->w[i] = window_for(list[i]);
        u64 x = 1;
        for (int d = 0; d < (1 << B->w[i]); d++) { B->pw[i][d] = x; x *= list[i]; }
    }
    B->nb = n;
    for (int w = MAXW; w >= 1; w--) {
        int c = B->ncls, k = 0;
        for (int j = 0; j < n; j++) if (B->w[j] == w) B->cls_j[c][k++] = j;
        if (k) { B->cls_w[c] = w; B->cls_n[c] = k; B->ncls++; }
    }
    return 0;
}

typedef struct {
    const pctx_t *c;
    const u64 *pw;
    u64 e;
    int j;
} lane_t;

/* LANES fixed-window exponentiations in lock step (wieferich.c:914-928). */
static void run_group(const lane_t *L, int w, u128 *r)
{
    const u64 mask = ((u64)1 << w) - 1;
    u128 x[LANES];
    u64 emax = 0;
    for (int g = 0; g < LANES; g++) emax |= L[g].e;
    int sh = ((64 - __builtin_clzll(emax) + w - 1) / w - 1) * w;
    for (int g = 0; g < LANES; g++) x[g] = msmall(L[g].c->one, L[g].pw[(L[g].e >> sh) & mask], L[g].c);
    for (sh -= w; sh >= 0; sh -= w) {
        for (int s = 0; s < w; s++)
            for (int g = 0; g < LANES; g++) x[g] = msqr(x[g], L[g].c);
        for (int g = 0; g < LANES; g++) x[g] = msmall(x[g], L[g].pw[(L[g].e >> sh) & mask], L[g].c);
    }
    for (int g = 0; g < LANES; g++) r[g] = redc(x[g], L[g].c);
}

/* b^(p-1) mod p^2 and q for one pair (wieferich.c:939-949; full path for p >= 2^63). */
u64 wfo_fermat_quotient(u64 p, u32 b, u64 *r_hi, u64 *r_lo)
{
    pctx_t c;
    pctx_init(&c, p);
    u128 r;
    if (c.full) {
        r = powmod_full(b, p - 1, &c);
    } else {
        bases_t B;
        bases_setup(&B, &b, 1);
        lane_t L[LANES];
        for (int g = 0; g < LANES; g++) { L[g].c = &c; L[g].pw = B.pw[0]; L[g].e = p - 1; L[g].j = 0; }
        u128 rr[LANES];
        run_group(L, B.w[0], rr);
        r = rr[0];
    }
    if (r_hi) *r_hi = (u64)(r >> 64);
    if (r_lo) *r_lo = (u64)r;
    return ((u64)r - 1) * c.pinv;
}

void wfo_pctx(u64 p, u64 *pinv, u64 *ninv, u64 *one_hi, u64 *one_lo)
{
    pctx_t c;
    pctx_init(&c, p);
    *pinv = c.pinv; *ninv = c.ninv;
    *one_hi = (u64)(c.one >> 64); *one_lo = (u64)c.one;
}

void wfo_msqr(u64 p, u64 x_hi, u64 x_lo, u64 *r_hi, u64 *r_lo)
{
    pctx_t c;
    pctx_init(&c, p);
    const u128 x = ((u128)x_hi << 64) | x_lo;
    const u128 r = c.full ? mmul_full(x, x, &c) : msqr(x, &c);
    *r_hi = (u64)(r >> 64); *r_lo = (u64)r;
}

void wfo_msmall(u64 p, u64 x_hi, u64 x_lo, u64 cst, u64 *r_hi, u64 *r_lo)
{
    pctx_t c;
    pctx_init(&c, p);
    const u128 x = ((u128)x_hi << 64) | x_lo;
    const u128 r = c.full ? mulmod_ref(x, cst, c.N) : msmall(x, cst, &c);
    *r_hi = (u64)(r >> 64); *r_lo = (u64)r;
}

void wfo_redc(u64 p, u64 x_hi, u64 x_lo, u64 *r_hi, u64 *r_lo)
{
    pctx_t c;
    pctx_init(&c, p);
    const u128 x = ((u128)x_hi << 64) | x_lo;
    const u128 r = c.full ? mmul_full(x, 1, &c) : redc(x, &c);
    *r_hi = (u64)(r >> 64); *r_lo = (u64)r;
}

/* ------------------------------------------------------------------ */
/* prime source: segmented Eratosthenes sieve over odd numbers          */
/* (replaces the external primesieve iterator, wieferich.c:1138-1141)   */
/* ------------------------------------------------------------------ */

static u32 *g_sp = NULL;        /* odd primes 3 .. g_sp_lim */
static u64 g_nsp = 0, g_sp_lim = 0;
/* Scans read the table under the read lock; growing it takes the write lock, so scans may
 * run concurrently (the tests run many at once from threads). */
static pthread_rwlock_t g_sp_rw = PTHREAD_RWLOCK_INITIALIZER;

static u64 isqrt64(u64 n)
{
    u64 r = (u64)sqrt((double)n);
    while (r && (u128)r * r > n) r--;
    while ((u128)(r + 1) * (r + 1) <= n) r++;
    return r;
}

/* Make g_sp hold every odd prime <= lim (lim <= 2^32) and return holding the table's read
 * lock: the caller reads g_sp, then calls sp_release.  The table only grows. */
static void sieving_primes(u64 lim)
{
    pthread_rwlock_rdlock(&g_sp_rw);
    if (lim <= g_sp_lim) return;
    pthread_rwlock_unlock(&g_sp_rw);
    pthread_rwlock_wrlock(&g_sp_rw);
    if (lim <= g_sp_lim) {                      /* another thread grew it meanwhile */
        pthread_rwlock_unlock(&g_sp_rw);
        pthread_rwlock_rdlock(&g_sp_rw);
        return;
    }
    if (lim < 1024) lim = 1024;
    /* bit i of the odd-only table = 2i + 1 */
    const u64 nbits = lim / 2 + 1;
    uint8_t *comp = calloc(nbits / 8 + 1, 1);
    for (u64 i = 1; (2 * i + 1) * (2 * i + 1) <= lim; i++) {
        if (comp[i >> 3] >> (i & 7) & 1) continue;
        const u64 q = 2 * i + 1;
        for (u64 k = (q * q) / 2; k < nbits; k += q) comp[k >> 3] |= (uint8_t)(1u << (k & 7));
    }
    u64 n = 0;
    for (u64 i = 1; i < nbits; i++) if (!(comp[i >> 3] >> (i & 7) & 1) && 2 * i + 1 <= lim) n++;
    u32 *sp = malloc(n * sizeof *sp);
    n = 0;
    for (u64 i = 1; i < nbits; i++) if (!(comp[i >> 3] >> (i & 7) & 1) && 2 * i + 1 <= lim) sp[n++] = (u32)(2 * i + 1);
    free(comp);
    free(g_sp);
    g_sp = sp;
    g_nsp = n;
    g_sp_lim = lim;
    pthread_rwlock_unlock(&g_sp_rw);
    pthread_rwlock_rdlock(&g_sp_rw);
}

static void sp_release(void) { pthread_rwlock_unlock(&g_sp_rw); }

#define SEG_ODDS (1u << 21)     /* odd numbers per sieve segment (256 KiB bitmap) */

/* Segmented sieve over the odd numbers of [lo, hi) (hi == 0 means 2^64).  Index i
 * stands for s0 + 2i, s0 = the first odd number >= max(lo, 3).  Every sieving prime
 * keeps the index of its next odd multiple, so a segment costs no divisions. */
typedef struct {
    u64 s0, nodd;               /* first odd number, odd numbers in the range */
    u64 nq;                     /* sieving primes in use (those <= sqrt(hi - 1)) */
    u64 *nxt;                   /* per sieving prime: index of the next multiple */
    u64 seg;                    /* current segment: indices [seg, seg + SEG_ODDS) */
    u64 *bits;                  /* bit i - seg set = composite (or 1) */
} oddsieve_t;

static bool oddsieve_init(oddsieve_t *S, u64 lo, u64 hi)
{
    memset(S, 0, sizeof *S);
    const u64 top = hi ? hi - 1 : UINT64_MAX;
    S->s0 = lo < 3 ? 3 : (lo | 1);
    if (S->s0 > top || S->s0 < 3) return false;
    S->nodd = (top - S->s0) / 2 + 1;
    const u64 lim = isqrt64(top);
    while (S->nq < g_nsp && g_sp[S->nq] <= lim) S->nq++;
    S->nxt = malloc((S->nq ? S->nq : 1) * sizeof *S->nxt);
    for (u64 k = 0; k < S->nq; k++) {
        const u64 q = g_sp[k], qq = q * q;
        u64 off;                                /* offset of the first odd multiple >= max(q^2, s0) */
        if (qq >= S->s0) off = qq - S->s0;
        else {
            off = (q - S->s0 % q) % q;
            if (off & 1) off += q;              /* s0 odd: an odd offset lands on an even multiple */
        }
        S->nxt[k] = off / 2;                    /* may lie beyond the range: never marked */
    }
    S->bits = malloc((SEG_ODDS / 64 + 1) * 8);
    S->seg = 0;
    return true;
}

/* Sieve the segment starting at index S->seg; returns its length in indices. */
static u64 oddsieve_segment(oddsieve_t *S)
{
    u64 len = S->nodd - S->seg;
    if (len > SEG_ODDS) len = SEG_ODDS;
    memset(S->bits, 0, ((len + 63) / 64) * 8);
    const u64 end = S->seg + len;
    for (u64 k = 0; k < S->nq; k++) {
        const u64 q = g_sp[k];
        u64 i = S->nxt[k];
        for (; i < end; i += q) S->bits[(i - S->seg) >> 6] |= 1ULL << ((i - S->seg) & 63);
        S->nxt[k] = i;
    }
    if (S->seg == 0 && S->s0 == 1) S->bits[0] |= 1;
    return len;
}

static void oddsieve_free(oddsieve_t *S)
{
    free(S->nxt);
    free(S->bits);
}

/* Primes in [lo, hi) by trial division + MR (verify_wieferich.py's approach). */
static u64 primes_mr(u64 lo, u64 hi, u64 *out, u64 cap)
{
    u64 n = 0;
    if (lo <= 2 && (hi == 0 || hi > 2)) { if (n < cap) out[n] = 2; n++; }
    u64 x = lo < 3 ? 3 : (lo | 1);
    for (; x >= 3 && (hi == 0 || x < hi); x += 2) {
        if (wfo_is_prime64(x)) { if (n < cap) out[n] = x; n++; }
        if (x >= UINT64_MAX - 1) break;
    }
    return n;
}

uint64_t wfo_primes(u64 lo, u64 hi, u64 *out, u64 cap, int method)
{
    if (hi != 0 && hi <= lo) return 0;
    if (method == 1) return primes_mr(lo, hi, out, cap);
    const u64 top = hi ? hi - 1 : UINT64_MAX;
    sieving_primes(isqrt64(top) + 1);
    u64 n = 0;
    if (lo <= 2 && (hi == 0 || hi > 2)) { if (n < cap) out[n] = 2; n++; }
    oddsieve_t S;
    if (oddsieve_init(&S, lo, hi)) {
        while (S.seg < S.nodd) {
            const u64 len = oddsieve_segment(&S);
            for (u64 i = 0; i < len; i++)
                if (!(S.bits[i >> 6] >> (i & 63) & 1)) { if (n < cap) out[n] = S.s0 + 2 * (S.seg + i); n++; }
            S.seg += len;
        }
        oddsieve_free(&S);
    }
    sp_release();
    return n;
}

/* ------------------------------------------------------------------ */
/* the scan (wieferich.c:1041-1177)                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    wfo_result r;
    wfo_event *ev;
    u64 nev, evcap;
    u64 *psp;
    u64 npsp, pspcap;
} acc_t;

static void push_event(acc_t *a, u32 type, u32 j, u64 p, u64 q, u128 r)
{
    if (a->nev == a->evcap) {
        a->evcap = a->evcap ? 2 * a->evcap : 64;
        a->ev = realloc(a->ev, a->evcap * sizeof *a->ev);
    }
    wfo_event *e = &a->ev[a->nev++];
    e->type = type; e->j = j; e->p = p; e->q = q;
    e->r_hi = (u64)(r >> 64); e->r_lo = (u64)r;
}

/* Account one tested pair: q = (r - 1) pinv, FLT check, checksum, hit / near
 * (finish_lanes, wieferich.c:1051-1075); also the sum of q. */
static void finish_pair(acc_t *a, const pctx_t *c, int j, u128 r, int64_t near_T)
{
    const u64 p = c->p;
    const u64 q = ((u64)r - 1) * c->pinv;
    a->r.tests++;
    if (q >= p) {
        a->r.flt++;
        push_event(a, 2, (u32)j, p, q, r);
        return;
    }
    a->r.cks[j] += wfo_mix64(q ^ (p * GOLD));
    a->r.sumq[j] += q;
    if (q == 0) {
        push_event(a, 0, (u32)j, p, q, r);
    } else if (near_T > 0) {
        const int64_t A = q <= p / 2 ? (int64_t)q : -(int64_t)(p - q);
        if ((A < 0 ? -A : A) <= near_T) {
            a->r.near++;
            push_event(a, 1, (u32)j, p, q, r);
        }
    }
}

/* Test np set-up primes against every base (process_batch, wieferich.c:1086-1113). */
static void process_batch(const bases_t *B, const pctx_t *ctx, int np, acc_t *a, int64_t near_T)
{
    lane_t L[LANES];
    u128 r[LANES];
    for (int c = 0; c < B->ncls; c++) {
        const int w = B->cls_w[c];
        int nl = 0;
        for (int i = 0; i < np; i++) {
            const u64 p = ctx[i].p;
            for (int k = 0; k < B->cls_n[c]; k++) {
                const int j = B->cls_j[c][k];
                if (p <= B->b[j] && B->b[j] % p == 0) continue;
                if (ctx[i].full) {
                    finish_pair(a, &ctx[i], j, powmod_full(B->b[j], p - 1, &ctx[i]), near_T);
                    continue;
                }
                L[nl].c = &ctx[i]; L[nl].pw = B->pw[j]; L[nl].e = p - 1; L[nl].j = j;
                if (++nl == LANES) {
                    run_group(L, w, r);
                    for (int g = 0; g < LANES; g++) finish_pair(a, L[g].c, L[g].j, r[g], near_T);
                    nl = 0;
                }
            }
        }
        if (nl) {
            for (int g = nl; g < LANES; g++) L[g] = L[nl - 1];
            run_group(L, w, r);
            for (int g = 0; g < nl; g++) finish_pair(a, L[g].c, L[g].j, r[g], near_T);
        }
    }
}

typedef struct {
    u64 lo, hi;                 /* hi == 0: 2^64 */
    u64 piece, npieces;
    atomic_ullong next;
    const bases_t *B;
    int64_t near_T;
    int want_psp;
} job_t;

typedef struct {
    job_t *job;
    acc_t acc;
} worker_t;

/* Scan [lo, hi) of one piece (scan_chunk, wieferich.c:1129-1177). */
static void scan_piece(job_t *J, acc_t *a, u64 lo, u64 hi)
{
    const bases_t *B = J->B;
    pctx_t ctx[BATCH];
    int np = 0;
    if (lo <= 2 && (hi == 0 || hi > 2)) {
        a->r.primes++;
        for (int j = 0; j < B->nb; j++) if (B->b[j] % 4 == 1) push_event(a, 0, (u32)j, 2, 0, 1);
    }
    oddsieve_t S;
    if (!oddsieve_init(&S, lo, hi)) return;
    while (S.seg < S.nodd) {
        const u64 len = oddsieve_segment(&S);
        for (u64 i = 0; i < len; i++) {
            const u64 n = S.s0 + 2 * (S.seg + i);
            if (S.bits[i >> 6] >> (i & 63) & 1) {
                if (J->want_psp && n > 1 && wfo_sprp2(n)) {
                    if (a->npsp == a->pspcap) {
                        a->pspcap = a->pspcap ? 2 * a->pspcap : 16;
                        a->psp = realloc(a->psp, a->pspcap * sizeof *a->psp);
                    }
                    a->psp[a->npsp++] = n;
                }
                continue;
            }
            a->r.primes++;
            pctx_init(&ctx[np], n);
            if (++np == BATCH) { process_batch(B, ctx, np, a, J->near_T); np = 0; }
        }
        S.seg += len;
    }
    if (np) process_batch(B, ctx, np, a, J->near_T);
    oddsieve_free(&S);
}

static void *worker_main(void *arg)
{
    worker_t *W = arg;
    job_t *J = W->job;
    for (;;) {
        const u64 k = atomic_fetch_add(&J->next, 1);
        if (k >= J->npieces) break;
        const u64 plo = J->lo + k * J->piece;
        u64 phi = plo + J->piece;
        if (k == J->npieces - 1) phi = J->hi;
        scan_piece(J, &W->acc, plo, phi);
    }
    return NULL;
}

static int cmp_event(const void *x, const void *y)
{
    const wfo_event *a = x, *b = y;
    if (a->p != b->p) return a->p < b->p ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    return (int)a->type - (int)b->type;
}

static int cmp_u64(const void *x, const void *y)
{
    const u64 a = *(const u64 *)x, b = *(const u64 *)y;
    return a < b ? -1 : a > b;
}

int wfo_scan(u64 lo, u64 hi, const u32 *bases, int nb, int64_t near_T, int threads,
             wfo_result *res, wfo_event *ev, u64 ev_cap, int want_psp, u64 *psp, u64 psp_cap)
{
    if ((hi != 0 && hi <= lo) || threads < 1 || !res) return -1;
    bases_t B;
    if (bases_setup(&B, bases, nb) != 0) return -1;
    const u64 top = hi ? hi - 1 : UINT64_MAX;
    sieving_primes(isqrt64(top) + 1);
    job_t J;
    memset(&J, 0, sizeof J);
    J.lo = lo; J.hi = hi; J.B = &B; J.near_T = near_T; J.want_psp = want_psp;
    const u64 span = top - lo + 1;              /* numbers in the range, 0 means 2^64 */
    u64 piece = span / ((u64)threads * 8) + 1;
    if (piece < (1u << 20)) piece = 1u << 20;
    piece = (piece + 1) & ~1ULL;                /* even, so every piece starts on the parity of lo */
    J.piece = piece;
    J.npieces = span ? (span - 1) / piece + 1 : (UINT64_MAX / piece) + 1;
    atomic_store(&J.next, 0);
    worker_t *W = calloc((size_t)threads, sizeof *W);
    pthread_t *th = calloc((size_t)threads, sizeof *th);
    for (int t = 0; t < threads; t++) {
        W[t].job = &J;
        if (pthread_create(&th[t], NULL, worker_main, &W[t]) != 0) worker_main(&W[t]), th[t] = 0;
    }
    for (int t = 0; t < threads; t++) if (th[t]) pthread_join(th[t], NULL);
    sp_release();
    memset(res, 0, sizeof *res);
    u64 nev = 0, npsp = 0;
    for (int t = 0; t < threads; t++) {
        wfo_result *r = &W[t].acc.r;
        res->primes += r->primes; res->tests += r->tests; res->near += r->near; res->flt += r->flt;
        for (int j = 0; j < nb; j++) { res->cks[j] += r->cks[j]; res->sumq[j] += r->sumq[j]; }
        nev += W[t].acc.nev;
        npsp += W[t].acc.npsp;
    }
    wfo_event *all = malloc((nev ? nev : 1) * sizeof *all);
    u64 *allp = malloc((npsp ? npsp : 1) * sizeof *allp);
    u64 k = 0, kp = 0;
    for (int t = 0; t < threads; t++) {
        if (W[t].acc.nev) memcpy(all + k, W[t].acc.ev, W[t].acc.nev * sizeof *all);   /* no copy from NULL */
        k += W[t].acc.nev;
        if (W[t].acc.npsp) memcpy(allp + kp, W[t].acc.psp, W[t].acc.npsp * sizeof *allp);
        kp += W[t].acc.npsp;
        free(W[t].acc.ev);
        free(W[t].acc.psp);
    }
    qsort(all, nev, sizeof *all, cmp_event);
    qsort(allp, npsp, sizeof *allp, cmp_u64);
    for (u64 i = 0; i < nev && i < ev_cap; i++) ev[i] = all[i];
    for (u64 i = 0; i < npsp && i < psp_cap; i++) psp[i] = allp[i];
    res->nev = nev;
    res->npsp = npsp;
    free(all);
    free(allp);
    free(W);
    free(th);
    return 0;
}
