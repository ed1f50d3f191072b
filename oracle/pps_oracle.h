/* pps_oracle.h -- CPU restatement of PPSieve's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (libppsd) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Ken-g6/PSieve-CUDA tree).
 */
#ifndef PPS_ORACLE_H
#define PPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ERR_INVALID_PARAM (-178) /* pps/app.h:67 */
#define ORC_ERR_FAILURE (-1)         /* main.c boinc_finish(EXIT_FAILURE) */

typedef struct {
  /* inputs as parsed from the command line (0 = not given) */
  uint64_t pmin, pmax;
  uint64_t kmin, kmax;
  uint32_t nmin, nmax;     /* nmax exclusive (pps/app.c:163) */
  uint32_t kstep, koffset; /* -M / -s (pps/app.c:595-604); 0 = defaults 2 / 1 */
  int32_t sign;            /* +1 Proth (default), -1 Riesel (-R), 0 = both */
  /* outputs of orc_derive */
  uint32_t nstep_ref; /* the reference's nstep after its clamps */
  uint32_t nstep;     /* window actually walked: min(nstep_ref, 32) */
  int32_t status;     /* 0 or ORC_ERR_* */
  char msg[160];      /* reason when status != 0 */
} orc_params;

typedef struct {
  uint64_t p;
  uint64_t k;
  uint32_t n;
  int32_t c;
} orc_hit;

int orc_derive(orc_params *pr);
uint64_t orc_isqrt(uint64_t x);
int orc_is_prime(uint64_t n);
int64_t orc_primes(uint64_t p0, uint64_t p1, uint64_t *out, uint64_t cap);
int orc_count_primes(uint64_t p0, uint64_t p1, int threads, uint64_t *count, uint64_t *sum);
int orc_count_primes_many(uint32_t n, const uint64_t *p0, const uint64_t *p1, int threads, uint64_t *count,
                          uint64_t *sum);
uint64_t orc_neg_inv64(uint64_t p);
uint64_t orc_inv2pow(uint64_t p, uint32_t n);
uint64_t orc_step(uint64_t K, uint64_t p, uint64_t ps, uint32_t s);
int orc_try_all_factors(uint64_t k, uint32_t n, int c);
int orc_test_factor(const orc_params *pr, uint64_t k, uint32_t n, int c);
int64_t orc_walk(const orc_params *pr, const uint64_t *P, uint64_t np,
                 orc_hit *out, uint64_t cap, int filter, int nthreads);
int64_t orc_run_range(const orc_params *pr, uint64_t p0, uint64_t p1,
                      orc_hit *out, uint64_t cap, int filter, int nthreads,
                      uint64_t *count, uint64_t *sum);
#define ORC_BAD_PRIME 1
#define ORC_BAD_DIVIDES 2
#define ORC_BAD_K 4
#define ORC_BAD_N 8
#define ORC_BAD_SIGN 16
#define ORC_KEPT 32
int64_t orc_check_hits(const orc_params *pr, const orc_hit *h, uint64_t n, int nthreads, uint8_t *flags);
int orc_format_hit(const orc_hit *h, char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
