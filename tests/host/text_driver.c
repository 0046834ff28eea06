/*
 * text_driver.c -- runs the wf host's text layer (mathpub_amd/host/wf_text.c) on its
 * arguments, one request per invocation, for tests/test_host_text.py.  The test builds it
 * twice: plain, and with -fsanitize=address,undefined to feed it malformed input.
 *
 *   number S                 -> "ok V" | "bad" | "overflow" | "negative"
 *   numbers                  -> the same for every line of stdin
 *   bases S                  -> "ok b1,b2,..." | "err MESSAGE"
 *   bases-list               -> the same for every line of stdin
 *   ckpt FILE                -> the parsed fields, one per line | "err MESSAGE"
 *   ckpt-rewrite FILE        -> the checkpoint as wt_ckpt_write writes it
 *   chunk LO HI PRIMES B:HEX...      -> the chunk-log line
 *   summary START POS PRIMES TESTS FLT NEAR_T NEAR FRONTIER SUMQ(0|1) B:HEX:SUMQHEX... [hit:B:P:CHUNK...]
 *   duration SECONDS         -> the status line's form
 */
#define _GNU_SOURCE
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wf_text.h"

static wt_u128 num(const char *s)
{
    wt_u128 v = 0;
    if (wt_number(s, &v) != WT_OK) {
        fprintf(stderr, "driver: bad argument %s\n", s);
        exit(2);
    }
    return v;
}

static void print_ckpt(const wt_ckpt *c)
{
    char end[24];
    wt_u128_str(end, sizeof end, c->end);
    printf("bases %s\nstart %" PRIu64 "\nchunk %" PRIu64 "\nend %s\nfrontier %" PRIu64 "\nprimes %" PRIu64
           "\ntests %" PRIu64 "\nnear %" PRIu64 "\nflt %" PRIu64 "\n",
           c->bases, c->start, c->chunk, end, c->frontier, c->primes, c->tests, c->near, c->flt);
    for (int j = 0; j < c->ncks; j++) printf("cks %u %016" PRIx64 "\n", c->cks_base[j], c->cks[j]);
    for (size_t i = 0; i < c->nhits; i++)
        printf("hit %u %" PRIu64 " %" PRIu64 "\n", c->hits[i].b, c->hits[i].p, c->hits[i].chunk);
}

/* One line of stdin per request (numbers / bases), for the fuzz corpus. */
static int lines(int bases)
{
    char *line = NULL;
    size_t cap = 0;
    ssize_t n;
    while ((n = getline(&line, &cap, stdin)) >= 0) {
        if (n && line[n - 1] == '\n') line[--n] = 0;
        if (bases) {
            uint32_t b[WT_MAXB];
            int nb = 0;
            char err[256], out[1024];
            if (wt_bases(line, b, &nb, err, sizeof err)) {
                wt_join_bases(out, sizeof out, b, nb);
                printf("ok %s\n", out);
            } else {
                printf("err %s\n", err);
            }
        } else {
            wt_u128 v = 0;
            const int rc = wt_number(line, &v);
            char s[24];
            wt_u128_str(s, sizeof s, v);
            if (rc == WT_OK) printf("ok %s\n", s);
            else puts(rc == WT_BAD ? "bad" : rc == WT_OVERFLOW ? "overflow" : "negative");
        }
    }
    free(line);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 2 && !strcmp(argv[1], "numbers")) return lines(0);
    if (argc == 2 && !strcmp(argv[1], "bases-list")) return lines(1);
    if (argc < 3) return 2;
    const char *cmd = argv[1];
    if (!strcmp(cmd, "number")) {
        wt_u128 v = 0;
        const int rc = wt_number(argv[2], &v);
        char s[24];
        wt_u128_str(s, sizeof s, v);
        if (rc == WT_OK) printf("ok %s\n", s);
        else puts(rc == WT_BAD ? "bad" : rc == WT_OVERFLOW ? "overflow" : "negative");
        return 0;
    }
    if (!strcmp(cmd, "bases")) {
        uint32_t b[WT_MAXB];
        int nb = 0;
        char err[256], out[1024];
        if (!wt_bases(argv[2], b, &nb, err, sizeof err)) {
            printf("err %s\n", err);
            return 0;
        }
        wt_join_bases(out, sizeof out, b, nb);
        printf("ok %s\n", out);
        return 0;
    }
    if (!strcmp(cmd, "ckpt") || !strcmp(cmd, "ckpt-rewrite")) {
        FILE *f = fopen(argv[2], "r");
        if (!f) return 2;
        wt_ckpt c;
        char err[2048];
        const int rc = wt_ckpt_read(f, argv[2], &c, err, sizeof err);
        fclose(f);
        if (rc) {
            printf("err %s\n", err);
            return 0;
        }
        if (!strcmp(cmd, "ckpt")) print_ckpt(&c);
        else if (!wt_ckpt_write(stdout, &c)) return 1;
        wt_ckpt_free(&c);
        return 0;
    }
    if (!strcmp(cmd, "chunk") && argc >= 5) {
        uint32_t b[WT_MAXB];
        uint64_t cks[WT_MAXB];
        int nb = 0;
        for (int i = 5; i < argc && nb < WT_MAXB; i++, nb++)
            if (sscanf(argv[i], "%" SCNu32 ":%" SCNx64, &b[nb], &cks[nb]) != 2) return 2;
        char line[64 + WT_MAXB * 32];
        wt_chunk_line(line, sizeof line, (uint64_t)num(argv[2]), num(argv[3]), (uint64_t)num(argv[4]), b, cks, nb);
        puts(line);
        return 0;
    }
    if (!strcmp(cmd, "summary") && argc >= 11) {
        uint32_t b[WT_MAXB];
        uint64_t cks[WT_MAXB], sumq[WT_MAXB];
        wt_hit hits[64];
        int nb = 0;
        size_t nh = 0;
        for (int i = 11; i < argc; i++) {
            if (!strncmp(argv[i], "hit:", 4)) {
                if (nh == 64 || sscanf(argv[i] + 4, "%" SCNu32 ":%" SCNu64 ":%" SCNu64, &hits[nh].b, &hits[nh].p,
                                       &hits[nh].chunk) != 3)
                    return 2;
                nh++;
            } else if (nb == WT_MAXB || sscanf(argv[i], "%" SCNu32 ":%" SCNx64 ":%" SCNx64, &b[nb], &cks[nb], &sumq[nb]) != 3) {
                return 2;
            } else {
                nb++;
            }
        }
        wt_summary(stdout, (uint64_t)num(argv[2]), num(argv[3]), (uint64_t)num(argv[4]), (uint64_t)num(argv[5]),
                   (uint64_t)num(argv[6]), (long long)num(argv[7]), (uint64_t)num(argv[8]), b, nb, cks,
                   atoi(argv[10]) ? sumq : NULL, hits, nh, (uint64_t)num(argv[9]));
        return 0;
    }
    if (!strcmp(cmd, "duration")) {
        char buf[32];
        wt_duration(buf, sizeof buf, strtod(argv[2], NULL));
        puts(buf);
        return 0;
    }
    return 2;
}
