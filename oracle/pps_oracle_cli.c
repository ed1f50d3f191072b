/* pps_oracle_cli.c -- command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (fixture generation and CPU timing).
 * Options follow the reference's names (main.c:375-398, pps/app.h:39-59):
 *   -p pmin -P pmax -k kmin -K kmax -n nmin -N nmax [-R] [--both]
 *   [-t threads] [--raw] [--count]
 * Value suffixes are not parsed here; pass plain integers.
 * Prints one "p | k*2^n+c" line per factor (pps/app.c:134), unsorted.
 */
#include <getopt.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pps_oracle.h"

int main(int argc, char **argv) {
  orc_params pr;
  int opt, threads = 1, raw = 0, count_only = 0, idx;
  static struct option lo[] = {{"both", no_argument, 0, 'b'},
                               {"raw", no_argument, 0, 'r'},
                               {"count", no_argument, 0, 'c'},
                               {0, 0, 0, 0}};
  memset(&pr, 0, sizeof(pr));
  pr.sign = 1;
  while ((opt = getopt_long(argc, argv, "p:P:k:K:n:N:Rt:", lo, &idx)) != -1) {
    switch (opt) {
      case 'p': pr.pmin = strtoull(optarg, 0, 0); break;
      case 'P': pr.pmax = strtoull(optarg, 0, 0); break;
      case 'k': pr.kmin = strtoull(optarg, 0, 0); break;
      case 'K': pr.kmax = strtoull(optarg, 0, 0); break;
      case 'n': pr.nmin = (uint32_t)strtoul(optarg, 0, 0); break;
      case 'N': pr.nmax = (uint32_t)strtoul(optarg, 0, 0); break;
      case 'R': pr.sign = -1; break;
      case 'b': pr.sign = 0; break;
      case 't': threads = atoi(optarg); break;
      case 'r': raw = 1; break;
      case 'c': count_only = 1; break;
      default: return 2;
    }
  }
  if (orc_derive(&pr) != 0) {
    fprintf(stderr, "invalid parameters: %s\n", pr.msg);
    return 1;
  }
  {
    uint64_t cnt = 0, sum = 0;
    struct timespec t0, t1;
    int64_t nh;
    orc_hit *hits;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    nh = orc_run_range(&pr, pr.pmin, pr.pmax, NULL, 0, !raw, threads, &cnt, &sum);
    hits = malloc(sizeof(orc_hit) * (size_t)(nh + 1));
    if (!count_only) nh = orc_run_range(&pr, pr.pmin, pr.pmax, hits, (uint64_t)nh, !raw, threads, &cnt, &sum);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (!count_only)
      for (int64_t i = 0; i < nh; i++) {
        char buf[96];
        orc_format_hit(&hits[i], buf, sizeof buf);
        fputs(buf, stdout);
      }
    fprintf(stderr, "nstep=%u (ref %u) nmin=%u nmax=%u kmin=%" PRIu64 " kmax=%" PRIu64 "\n", pr.nstep,
            pr.nstep_ref, pr.nmin, pr.nmax, pr.kmin, pr.kmax);
    fprintf(stderr, "count=%" PRIu64 ",sum=0x%016" PRIx64 " hits=%" PRId64 " seconds=%.3f\n", cnt, sum, nh,
            (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
    free(hits);
  }
  return 0;
}
