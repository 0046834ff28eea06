/*
 * oracle_driver.c -- runs the CPU oracle's scan (oracle/wf_oracle.c) on one range from the
 * command line and prints its totals and events, for tests/test_host_text.py, which builds
 * it with -fsanitize=address,undefined and compares its output with the plain library's.
 *
 *   oracle_driver LO HI NEAR_T PSP THREADS B1 [B2 ...]      (HI = 0 means 2^64)
 *   oracle_driver q P1 [P2 ...]                             (quotients for bases 2 and 3)
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wf_oracle.h"

int main(int argc, char **argv)
{
    if (argc >= 3 && !strcmp(argv[1], "q")) {      /* quotients of single p for bases 2 and 3 */
        for (int i = 2; i < argc; i++) {
            const uint64_t p = strtoull(argv[i], NULL, 10);
            for (uint32_t b = 2; b <= 3; b++) {
                uint64_t hi, lo;
                const uint64_t q = wfo_fermat_quotient(p, b, &hi, &lo);
                printf("%" PRIu64 " %u %" PRIu64 " %016" PRIx64 "%016" PRIx64 "\n", p, b, q, hi, lo);
            }
        }
        return 0;
    }
    if (argc < 7) return 2;
    const uint64_t lo = strtoull(argv[1], NULL, 10), hi = strtoull(argv[2], NULL, 10);
    const int64_t nearT = strtoll(argv[3], NULL, 10);
    const int psp = atoi(argv[4]), threads = atoi(argv[5]);
    uint32_t b[WFO_MAXB];
    int nb = 0;
    for (int i = 6; i < argc && nb < WFO_MAXB; i++) b[nb++] = (uint32_t)strtoul(argv[i], NULL, 10);
    enum { EVCAP = 1 << 12, PSPCAP = 1 << 12 };
    static wfo_event ev[EVCAP];
    static uint64_t ps[PSPCAP];
    wfo_result r;
    if (wfo_scan(lo, hi, b, nb, nearT, threads, &r, ev, EVCAP, psp, ps, PSPCAP) != 0) return 1;
    printf("primes %" PRIu64 " tests %" PRIu64 " near %" PRIu64 " flt %" PRIu64 " nev %" PRIu64 " npsp %" PRIu64 "\n",
           r.primes, r.tests, r.near, r.flt, r.nev, r.npsp);
    for (int j = 0; j < nb; j++) printf("base %u cks %016" PRIx64 " sumq %016" PRIx64 "\n", b[j], r.cks[j], r.sumq[j]);
    for (uint64_t k = 0; k < r.nev && k < EVCAP; k++)
        printf("ev %u %u %" PRIu64 " %" PRIu64 "\n", ev[k].type, ev[k].j, ev[k].p, ev[k].q);
    for (uint64_t k = 0; k < r.npsp && k < PSPCAP; k++) printf("psp %" PRIu64 "\n", ps[k]);
    return 0;
}
