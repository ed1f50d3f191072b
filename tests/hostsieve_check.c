/* hostsieve_check.c -- prints the prime count and sum (mod 2^64) of each
 * range [p0, p1) given on the command line, as the host sieve of
 * `ppsieve --host-primes` supplies them: "p0 p1 count sum_hex order" per
 * line.  A first argument -tN sieves with N threads. */
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "pps_hostsieve.h"

struct acc {
  uint64_t count, sum, last;
  int ordered;
};

static void take(void *arg, const uint64_t *p, size_t n) {
  struct acc *a = arg;
  for (size_t i = 0; i < n; i++) {
    if (p[i] <= a->last) a->ordered = 0;
    a->last = p[i];
    a->count++;
    a->sum += p[i];
  }
}

int main(int argc, char **argv) {
  uint64_t top = 0;
  int nthreads = 1;
  if (argc > 1 && argv[1][0] == '-' && argv[1][1] == 't') {
    nthreads = atoi(argv[1] + 2);
    argv++;
    argc--;
  }
  for (int i = 2; i < argc; i += 2) {
    const uint64_t p1 = strtoull(argv[i], NULL, 0);
    if (p1 > top) top = p1;
  }
  uint32_t lim = 3;
  while ((uint64_t)lim * lim < top) lim++;
  unsigned char *comp = calloc(lim / 2 + 1, 1);
  uint32_t *q = malloc(sizeof(uint32_t) * (lim / 2 + 1)), nq = 0;
  if (!comp || !q) return 2;
  for (uint64_t i = 3; i <= lim; i += 2)
    if (!comp[i / 2]) {
      q[nq++] = (uint32_t)i;
      for (uint64_t m = i * i; m <= lim; m += 2 * i) comp[m / 2] = 1;
    }
  pps_hostsieve *h = pps_hostsieve_new(q, nq, nthreads);
  if (!h) return 2;
  for (int i = 1; i + 1 < argc; i += 2) {
    const uint64_t p0 = strtoull(argv[i], NULL, 0), p1 = strtoull(argv[i + 1], NULL, 0);
    struct acc a = {0, 0, 0, 1};
    if (pps_hostsieve_range(h, p0, p1, take, &a)) return 2;
    printf("%" PRIu64 " %" PRIu64 " %" PRIu64 " 0x%" PRIx64 " %s\n", p0, p1, a.count, a.sum,
           a.ordered ? "ascending" : "UNORDERED");
  }
  pps_hostsieve_free(h);
  free(q);
  free(comp);
  return 0;
}
