/*
 * wf_text.h -- the text formats of the wf host program, with no engine dependency.
 *
 * The formats are the drop-in interface of A001220/wieferich.c and are kept byte for byte:
 *   numbers      TERM or TERM+TERM / TERM-TERM, a term being decimal digits, B^E, or M e E
 *                with one optional decimal point in M (wieferich.c:329-416);
 *   base lists   comma-separated numbers, LO-HI ranges, "oeis", "known" (wieferich.c:816-849);
 *   chunk log    "C lo hi primes b:cks ..." (wieferich.c:1166-1172);
 *   checkpoint   "wieferich checkpoint 1" / bases / start / chunk / end / frontier / primes /
 *                tests / near / flt / cks b hex ... / hit b p chunk ... / end
 *                (wieferich.c:1211-1240);
 *   SUMMARY      wieferich.c:1505-1523, plus "sumq base b hex" lines on request.
 * Every function reports errors by its return value; none exits or prints.  Values may
 * reach 2^64 (the end of the top-of-word range), so they are carried in 128 bits.
 */
#ifndef WF_TEXT_H
#define WF_TEXT_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

typedef unsigned __int128 wt_u128;

#define WT_TWO64 (((wt_u128)1) << 64)
#define WT_MAXB 64              /* bases per run (WF_MAXB of wf.h) */
#define WT_TERM_MAX 63          /* characters in one term of a number */
#define WT_BASES_MAX 1023       /* characters in a base list */

/* Results of wt_number. */
enum { WT_OK = 0, WT_BAD, WT_OVERFLOW, WT_NEGATIVE };

/* Parse a command-line number (at most 2^64).  WT_BAD for text the grammar rejects, a
 * term above 2^64 or a non-integer M e E; WT_OVERFLOW for a sum above 2^64; WT_NEGATIVE
 * for a difference below zero. */
int wt_number(const char *s, wt_u128 *out);

/* Parse a base list into ascending distinct bases.  On failure returns false and writes a
 * message (without trailing newline) into err. */
bool wt_bases(const char *s, uint32_t *out, int *nb, char *err, size_t errlen);

/* The reference's base lists: the 17 bases with an OEIS sequence of their own ("oeis",
 * the scan's default) and the bases 2..30 that are not perfect powers ("known", the check
 * command's default). */
extern const uint32_t WT_OEIS[17];
extern const uint32_t WT_KNOWN[23];

/* The OEIS A-number of base b's Wieferich-prime sequence, or NULL. */
const char *wt_anum(uint32_t b);

/* "b1,b2,..." (the checkpoint's and the status line's form). */
void wt_join_bases(char *buf, size_t n, const uint32_t *b, int nb);

/* Decimal text of v <= 2^64. */
void wt_u128_str(char *buf, size_t n, wt_u128 v);

/* The chunk-log line of [lo, hi) without its newline. */
void wt_chunk_line(char *buf, size_t n, uint64_t lo, wt_u128 hi, uint64_t primes, const uint32_t *b,
                   const uint64_t *cks, int nb);

/* A duration as the status line shows it: "12.3s", "4m05s", "2h03m04s", "3d04h05m", or "?"
 * for a negative, NaN or absurd value. */
void wt_duration(char *buf, size_t n, double s);

/* A solution found by a scan: base, prime, index of its chunk. */
typedef struct { uint32_t b; uint64_t p, chunk; } wt_hit;

/* The contents of a checkpoint file. */
typedef struct {
    char bases[WT_BASES_MAX + 1];
    uint64_t start, chunk, frontier, primes, tests, near, flt;
    wt_u128 end;
    int ncks;
    uint32_t cks_base[WT_MAXB];
    uint64_t cks[WT_MAXB];
    size_t nhits;
    wt_hit *hits;               /* malloc'd (wt_ckpt_free) */
} wt_ckpt;

/* Read a checkpoint.  Returns 0, or on failure one of the WT_CK_* codes with a message
 * naming `path` in err: not a checkpoint, truncated (no "end" line), or a malformed line. */
enum { WT_CK_NOTCK = 1, WT_CK_TRUNC, WT_CK_LINE, WT_CK_NOMEM };
int wt_ckpt_read(FILE *f, const char *path, wt_ckpt *c, char *err, size_t errlen);
void wt_ckpt_free(wt_ckpt *c);

/* Write a checkpoint; the hits written are those below the frontier.  Returns false if a
 * write failed. */
bool wt_ckpt_write(FILE *f, const wt_ckpt *c);

/* The SUMMARY block: [start, pos), totals, near count when near_T > 0, then per base its
 * A-number, checksum and the solutions below the frontier, and with sumq != NULL the sum
 * of the quotients. */
void wt_summary(FILE *f, uint64_t start, wt_u128 pos, uint64_t primes, uint64_t tests, uint64_t flt,
                long long near_T, uint64_t near, const uint32_t *b, int nb, const uint64_t *cks,
                const uint64_t *sumq, const wt_hit *hits, size_t nhits, uint64_t frontier);

#endif
