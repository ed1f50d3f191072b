/* pps_oracle.c -- CPU restatement of PPSieve's hot path (Proth/Riesel n-walk).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle that the HIP path is
 * checked against, and the CPU baseline that bench.py times.  The shipped
 * library never links it.
 *
 * What it restates (reference paths relative to the Ken-g6/PSieve-CUDA tree):
 *   parameter rules      main.c:916-932, pps/app.c:655-712,
 *                        pps/appcu.cu:155-209, pps/appcu.cu:236-239
 *   prime supply         sieve.c:168-608 + main.c:750-777 (primes in [pmin,pmax))
 *   per-prime walk       pps/app.c:1037-1121 (test_one_p)
 *   Montgomery helpers   pps/app.c:895-1035
 *   factor filter        pps/app.c:150-186 (test_factor),
 *                        pps/factor_proth.c:26-42, 220-253 (try_all_factors)
 *   output line          pps/app.c:134
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_derive           reference error rules, re-derived case by case
 *   orc_primes           sympy.primerange / isprime, count+sum, both code paths
 *   orc_neg_inv64        p * ps == -1 (mod 2^64)
 *   orc_inv2pow/orc_step Python pow(2, -n, p); the pps4test.ods halving table
 *   orc_walk             brute force over every (k, n) on tiny ranges; the
 *                        pps/oldsort.txt raw hits and their filtered subset
 *   orc_try_all_factors  trial division of k*2^n+c by every odd prime < 32768,
 *                        and the compiled reference pps/factor_proth.c driver
 */
#define _GNU_SOURCE
#include "pps_oracle.h"

#include <inttypes.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* parameter rules                                                      */
/* ------------------------------------------------------------------ */

/* lg2 as in pps/app.c:642-647: floor(log2(v)), 0 for v <= 1 */
static int lg2_u64(uint64_t v) {
  int r = 0;
  while (v >>= 1) r++;
  return r;
}

static int fail(orc_params *pr, int code, const char *msg) {
  pr->status = code;
  snprintf(pr->msg, sizeof(pr->msg), "%s", msg);
  return code;
}

/* Restates main.c:916-932 (pmin/pmax defaults), pps/app.c:655-712
 * (validation, nmin default, k alignment, nstep), pps/appcu.cu:155-209
 * (nstep clamps, 2^(64-nstep) <= pmin, nmin >= 64) and pps/appcu.cu:236-239
 * (pmin > 2^32). */
int orc_derive(orc_params *pr) {
  const uint64_t PMAX_MAX = UINT64_C(1) << 62;
  uint64_t b0, b1;
  uint32_t s;

  pr->status = 0;
  pr->msg[0] = 0;
  if (pr->kstep == 0) {
    pr->kstep = 2;
    pr->koffset = 1;
  }
  if (pr->pmin < 3) pr->pmin = 3;
  if (pr->pmax > PMAX_MAX) pr->pmax = PMAX_MAX;
  if (pr->pmin >= pr->pmax) {
    if (pr->pmax == 0 && pr->pmin < PMAX_MAX - 1000000000u)
      pr->pmax = pr->pmin + 1000000000u;
    else
      return fail(pr, ORC_ERR_FAILURE, "pmax must be greater than pmin");
  }
  if (pr->kmin == 0 || pr->kmax == 0)
    return fail(pr, ORC_ERR_INVALID_PARAM, "Please specify an input file or all of kmin, kmax, and nmax");
  if (pr->kmin > pr->kmax) return fail(pr, ORC_ERR_INVALID_PARAM, "kmin <= kmax is required");
  if (pr->kmax >= pr->pmin) return fail(pr, ORC_ERR_INVALID_PARAM, "kmax < pmin is required");
  if (pr->kmax - pr->kmin >= (UINT64_C(3) << 36))
    return fail(pr, ORC_ERR_INVALID_PARAM, "kmax-kmin < 3*2^36 is required");
  /* -n / -N parse as unsigned at most 2^31-1 (pps/app.c:579-580) */
  if (pr->nmin > 0x7fffffffu || pr->nmax > 0x7fffffffu)
    return fail(pr, ORC_ERR_INVALID_PARAM, "nmin and nmax must be at most 2^31-1");
  if (pr->nmin == 0) pr->nmin = (uint32_t)(2 * lg2_u64(pr->pmin) - lg2_u64(pr->kmax) - 1);
  if (pr->nmax == 0) pr->nmax = pr->nmin;
  if (pr->nmin > pr->nmax) return fail(pr, ORC_ERR_INVALID_PARAM, "nmin <= nmax is required");

  b0 = pr->kmin / pr->kstep;
  b1 = pr->kmax / pr->kstep;
  pr->kmin = b0 * pr->kstep + pr->koffset;
  pr->kmax = b1 * pr->kstep + pr->koffset;

  for (s = 1; (pr->kmax << s) < pr->pmin; s++)
    ;
  if (s > pr->nmax - pr->nmin + 1) s = pr->nmax - pr->nmin + 1;
  if ((UINT64_C(1) << (64 - s)) > pr->pmin)
    return fail(pr, ORC_ERR_INVALID_PARAM, "pmin is not large enough (or nmax is close to nmin)");
  if (s >= 32 && s < 48 && (UINT64_C(1) << 32) <= pr->pmin) s = 32;
  if (lg2_u64(pr->nmin) < 6)
    return fail(pr, ORC_ERR_INVALID_PARAM, "nmin too small (must be at least 64)");
  if (pr->pmin <= (UINT64_C(1) << 32)) return fail(pr, ORC_ERR_INVALID_PARAM, "PMin is too small, <= 2^32");
  pr->nstep_ref = s;
  /* s >= 48 is left unclamped by the reference (appcu.cu:182), and its CPU
   * re-walk truncates k0*Ps to 32x32 bits there (app.c:871-874, 1087).  The
   * factor set does not depend on the window, so the walk uses <= 32. */
  pr->nstep = s > 32 ? 32 : s;
  return 0;
}

/* ------------------------------------------------------------------ */
/* prime supply                                                         */
/* ------------------------------------------------------------------ */

uint64_t orc_isqrt(uint64_t x) {
  uint64_t r = 0, bit = UINT64_C(1) << 62;
  while (bit > x) bit >>= 2;
  while (bit) {
    if (x >= r + bit) {
      x -= r + bit;
      r = (r >> 1) + bit;
    } else
      r >>= 1;
    bit >>= 2;
  }
  return r;
}

static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

static uint64_t powmod64(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod64(r, b, m);
    b = mulmod64(b, b, m);
    e >>= 1;
  }
  return r;
}

/* Deterministic Miller-Rabin for 64-bit n (bases = first 12 primes). */
int orc_is_prime(uint64_t n) {
  static const uint32_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  uint64_t d;
  int r, i, j;
  if (n < 2) return 0;
  for (i = 0; i < 12; i++) {
    if (n == bases[i]) return 1;
    if (n % bases[i] == 0) return 0;
  }
  d = n - 1;
  r = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    r++;
  }
  for (i = 0; i < 12; i++) {
    uint64_t x = powmod64(bases[i], d, n);
    if (x == 1 || x == n - 1) continue;
    for (j = 1; j < r; j++) {
      x = mulmod64(x, x, n);
      if (x == n - 1) break;
    }
    if (j == r) return 0;
  }
  return 1;
}

/* x mod q for q < 2^32 without a 64-bit division (the sieves take one per
 * sieving prime per segment): a quotient estimate in double precision is
 * within 2^10 of the true one, the remainder it leaves fits 2^42, and a
 * second estimate on that is off by at most one. */
static inline uint64_t mod_u32(uint64_t x, uint32_t q) {
  const double dq = (double)q;
  const uint64_t q1 = (uint64_t)((double)x / dq);
  int64_t r = (int64_t)(x - q1 * (uint64_t)q);
  r -= (int64_t)((double)r / dq) * (int64_t)q;
  if (r < 0) r += q;
  if (r >= (int64_t)q) r -= q;
  return (uint64_t)r;
}

/* odd primes <= lim, ascending; returns count (caller frees *out) */
static uint32_t small_sieve(uint32_t lim, uint32_t **out) {
  uint32_t nodd = lim >= 3 ? (lim - 1) / 2 : 0; /* odd numbers 3..lim */
  uint8_t *comp = calloc(nodd + 1, 1);
  uint32_t i, cnt = 0, *res;
  for (i = 0; i < nodd; i++) {
    uint64_t q = 2 * (uint64_t)i + 3;
    if (comp[i]) continue;
    if (q * q > lim) continue;
    for (uint64_t m = q * q; m <= lim; m += 2 * q) comp[(m - 3) / 2] = 1;
  }
  for (i = 0; i < nodd; i++) cnt += !comp[i];
  res = malloc(sizeof(uint32_t) * (cnt + 1));
  cnt = 0;
  for (i = 0; i < nodd; i++)
    if (!comp[i]) res[cnt++] = 2 * i + 3;
  free(comp);
  *out = res;
  return cnt;
}

/* All primes p with p0 <= p < p1, ascending.  Returns the count; writes at
 * most cap of them.  Candidates are the odd p in [p0|1, p1) as in
 * sieve.c:729-760; composites are removed with every odd prime
 * q <= isqrt(p1-1) (the reference uses (unsigned)sqrt((double)pmax),
 * main.c:971-972).  A window that is narrow for its height is instead
 * presieved with the odd primes below 2^16 and its survivors tested by
 * Miller-Rabin, which yields the same set; the choice is by estimated
 * cost (a remainder per sieving prime per segment against about 20 us per
 * prime tested). */
static pthread_mutex_t base_mu = PTHREAD_MUTEX_INITIALIZER;
static uint32_t *base_q, base_n, base_lim;
static const uint32_t *base_primes(uint32_t lim, uint32_t *n);

static int64_t primes_mr(uint64_t lo, uint64_t p1, uint64_t *out, uint64_t cap, uint64_t n) {
  uint32_t nq;
  const uint32_t *qs = base_primes(65536, &nq);
  const uint64_t SEG = UINT64_C(1) << 18;
  uint8_t *seg = malloc(SEG);
  for (uint64_t base = lo; base < p1; base += 2 * SEG) {
    uint64_t cnt = (p1 - base + 1) / 2;
    if (cnt > SEG) cnt = SEG;
    memset(seg, 0, cnt);
    for (uint32_t qi = 0; qi < nq; qi++) {
      const uint64_t q = qs[qi];
      if (q >= 65536 || q * q >= base + 2 * cnt) break;
      uint64_t first = q * q;
      if (first < base) {
        const uint64_t r = mod_u32(base, (uint32_t)q);
        first = r ? base + (q - r) : base;
        if ((first & 1) == 0) first += q;
      }
      for (uint64_t m = first; m < base + 2 * cnt; m += 2 * q) seg[(m - base) / 2] = 1;
    }
    for (uint64_t j = 0; j < cnt; j++) {
      const uint64_t x = base + 2 * j;
      if (!seg[j] && x < p1 && orc_is_prime(x)) {
        if (n < cap) out[n] = x;
        n++;
      }
    }
  }
  free(seg);
  return (int64_t)n;
}

int64_t orc_primes(uint64_t p0, uint64_t p1, uint64_t *out, uint64_t cap) {
  uint64_t lim, n = 0, lo;
  if (p1 <= p0) return 0;
  if (p0 <= 2 && p1 > 2) {
    if (n < cap) out[n] = 2;
    n++;
  }
  lo = p0 | 1;
  if (lo < 3) lo = 3;
  if (lo >= p1) return (int64_t)n;
  lim = orc_isqrt(p1 - 1);
  {
    const double odds = (double)(p1 - lo) / 2, nq = (double)lim / (log((double)lim + 2) + 1);
    const double sieve_cost = (floor(odds / (1 << 20)) + 1) * nq * 6e-9 + (base_lim >= lim ? 0 : (double)lim * 2e-9);
    const double mr_cost = 2 * odds / log((double)p1) * 20e-6 + odds * 0.15 * 1.6e-6;
    if (lim > 0x7fffffffu || mr_cost < sieve_cost) return primes_mr(lo, p1, out, cap, n);
  }
  {
    uint32_t nq;
    const uint32_t *qs = base_primes((uint32_t)lim, &nq);
    const uint64_t SEG = UINT64_C(1) << 20; /* odd numbers per segment */
    uint8_t *seg = malloc(SEG);
    for (uint64_t base = lo; base < p1; base += 2 * SEG) {
      uint64_t cnt = (p1 - base + 1) / 2;
      if (cnt > SEG) cnt = SEG;
      memset(seg, 0, cnt);
      for (uint32_t qi = 0; qi < nq; qi++) {
        uint64_t q = qs[qi], first, r;
        if (q * q >= base + 2 * cnt) break;
        first = q * q;
        if (first < base) {
          r = mod_u32(base, (uint32_t)q);
          first = r ? base + (q - r) : base;
          if ((first & 1) == 0) first += q;
        }
        for (uint64_t m = first; m < base + 2 * cnt; m += 2 * q) seg[(m - base) / 2] = 1;
      }
      for (uint64_t j = 0; j < cnt; j++)
        if (!seg[j]) {
          uint64_t x = base + 2 * j;
          if (x < p1 && x > 1) {
            if (n < cap) out[n] = x;
            n++;
          }
        }
    }
    free(seg);
  }
  return (int64_t)n;
}

/* Count and sum (mod 2^64) of the primes in each of several ranges
 * [p0[i], p1[i]), p0 >= 3, for ranges too large to list: every range is cut
 * into segments of as many odd numbers as there are sieving primes (2^19 to
 * 2^27), and `threads` threads take (range,
 * segment) jobs in turn, sieving each with every odd prime <= isqrt of the
 * range's end.  The same set as orc_primes (the reference's count= and sum=,
 * main.c:1232).  The odd primes up to the largest isqrt seen are kept for
 * later calls (they take seconds to make near 2^62). */

static const uint32_t *base_primes(uint32_t lim, uint32_t *n) {
  pthread_mutex_lock(&base_mu);
  if (!base_q || lim > base_lim) {
    uint32_t *q;
    uint32_t cnt = small_sieve(lim, &q);
    free(base_q);
    base_q = q;
    base_n = cnt;
    base_lim = lim;
  }
  *n = base_n;
  pthread_mutex_unlock(&base_mu);
  return base_q;
}

struct count_job {
  const uint64_t *lo, *p1;
  uint32_t nranges;
  uint64_t seg;
  uint32_t nq;
  const uint32_t *qs;
  uint32_t cur;      /* range of the next job */
  uint64_t next;     /* start of the next job in that range */
  uint64_t *count, *sum;
  pthread_mutex_t mu;
};

static void *count_worker(void *arg) {
  struct count_job *J = arg;
  uint8_t *buf = malloc(J->seg);
  for (;;) {
    uint32_t ri;
    uint64_t base;
    pthread_mutex_lock(&J->mu);
    while (J->cur < J->nranges && J->next >= J->p1[J->cur]) {
      if (++J->cur < J->nranges) J->next = J->lo[J->cur];
    }
    ri = J->cur;
    base = J->next;
    J->next += 2 * J->seg;
    pthread_mutex_unlock(&J->mu);
    if (ri >= J->nranges) break;
    const uint64_t p1 = J->p1[ri];
    uint64_t cnt = (p1 - base + 1) / 2, count = 0, sum = 0;
    if (cnt > J->seg) cnt = J->seg;
    const uint64_t end = base + 2 * cnt;
    memset(buf, 0, cnt);
    for (uint32_t qi = 0; qi < J->nq; qi++) {
      const uint64_t q = J->qs[qi];
      if (q * q >= end) break;
      uint64_t first = q * q;
      if (first < base) {
        const uint64_t r = mod_u32(base, (uint32_t)q);
        first = r ? base + (q - r) : base;
        if ((first & 1) == 0) first += q;
      }
      for (uint64_t m = first; m < end; m += 2 * q) buf[(m - base) / 2] = 1;
    }
    for (uint64_t j = 0; j < cnt; j++)
      if (!buf[j]) {
        count++;
        sum += base + 2 * j;
      }
    pthread_mutex_lock(&J->mu);
    J->count[ri] += count;
    J->sum[ri] += sum;
    pthread_mutex_unlock(&J->mu);
  }
  free(buf);
  return NULL;
}

int orc_count_primes_many(uint32_t n, const uint64_t *p0, const uint64_t *p1, int threads, uint64_t *count,
                          uint64_t *sum) {
  struct count_job J;
  pthread_t th[64];
  uint64_t *lo, top = 0;
  if (threads < 1 || threads > 64) return -1;
  for (uint32_t i = 0; i < n; i++) {
    if (p0[i] < 3 || p1[i] <= p0[i]) return -1;
    if (p1[i] > top) top = p1[i];
  }
  if (n == 0) return 0;
  if (orc_isqrt(top - 1) > (UINT64_C(1) << 32) - 1) return -1;
  lo = malloc(sizeof(uint64_t) * n);
  for (uint32_t i = 0; i < n; i++) {
    lo[i] = p0[i] | 1;
    count[i] = sum[i] = 0;
  }
  memset(&J, 0, sizeof J);
  J.lo = lo;
  J.p1 = p1;
  J.nranges = n;
  J.qs = base_primes((uint32_t)orc_isqrt(top - 1), &J.nq);
  /* segments small enough to stay in cache while they are marked, large
   * enough that the one remainder per sieving prime per segment is noise */
  J.seg = (uint64_t)J.nq;
  if (J.seg < (UINT64_C(1) << 19)) J.seg = UINT64_C(1) << 19;
  if (J.seg > (UINT64_C(1) << 27)) J.seg = UINT64_C(1) << 27;
  J.cur = 0;
  J.next = lo[0];
  J.count = count;
  J.sum = sum;
  pthread_mutex_init(&J.mu, NULL);
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, count_worker, &J);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&J.mu);
  free(lo);
  return 0;
}

int orc_count_primes(uint64_t p0, uint64_t p1, int threads, uint64_t *count, uint64_t *sum) {
  return orc_count_primes_many(1, &p0, &p1, threads, count, sum);
}

/* ------------------------------------------------------------------ */
/* 64-bit Montgomery arithmetic, R = 2^64                               */
/* ------------------------------------------------------------------ */

/* -p^-1 mod 2^64 by Newton's iteration; restates invmod2pow_ul,
 * pps/app.c:895-914 (the reference negates at app.c:1066). */
uint64_t orc_neg_inv64(uint64_t p) {
  uint64_t x = p; /* p*p == 1 mod 8: 3 correct bits */
  for (int i = 0; i < 5; i++) x *= 2 - p * x;
  return (uint64_t)0 - x;
}

/* REDC(T) = T * 2^-64 mod p for T < p * 2^64 (mulmod_REDC, app.c:916-940) */
static inline uint64_t redc(u128 T, uint64_t p, uint64_t ps) {
  uint64_t m = (uint64_t)T * ps;
  u128 t = (T + (u128)m * p) >> 64;
  return t >= p ? (uint64_t)(t - p) : (uint64_t)t;
}

/* 2^-n mod p.  Restates invpowmod_REDClr + mod_REDC (app.c:1014-1035) and
 * the r0/bbits set-up (appcu.cu:193-204): start from the Montgomery form of
 * 2^-t, t = top 6 bits of n, then square and halve left to right. */
uint64_t orc_inv2pow(uint64_t p, uint32_t n) {
  uint64_t ps, r;
  int bb, b;
  if (n < 64) { /* the reference needs n >= 64; small n by plain halving */
    uint64_t x = 1 % p;
    for (uint32_t i = 0; i < n; i++) x = (x & 1) ? (uint64_t)(((u128)x + p) >> 1) : x >> 1;
    return x;
  }
  ps = orc_neg_inv64(p);
  bb = lg2_u64(n);
  r = (UINT64_C(1) << (64 - (n >> (bb - 5)))) % p;
  for (b = bb - 6; b >= 0; b--) {
    r = redc((u128)r * r, p, ps);
    if ((n >> b) & 1) r = (r & 1) ? (uint64_t)(((u128)r + p) >> 1) : r >> 1;
  }
  return redc(r, p, ps);
}

/* K * 2^-s mod p as REDC(K << (64-s)); restates shiftmod_REDC,
 * app.c:963-985.  Requires 1 <= s <= 63, K < p. */
uint64_t orc_step(uint64_t K, uint64_t p, uint64_t ps, uint32_t s) {
  return redc((u128)K << (64 - s), p, ps);
}

/* ------------------------------------------------------------------ */
/* small-prime filter                                                   */
/* ------------------------------------------------------------------ */

static uint32_t *g_small_primes; /* odd primes 11 <= q < 32768 */
static uint32_t g_num_small;
static pthread_once_t g_small_once = PTHREAD_ONCE_INIT;

/* The table of pps/factor_proth.c:26-42 (sieve_small_primes(11), app.c:739),
 * cut at 32768 as try_all_factors does (factor_proth.c:249). */
static void init_small_primes(void) {
  uint32_t *all, na = small_sieve(32767, &all), j = 0;
  g_small_primes = malloc(sizeof(uint32_t) * na);
  for (uint32_t i = 0; i < na; i++)
    if (all[i] >= 11) g_small_primes[j++] = all[i];
  g_num_small = j;
  free(all);
}

/* Smallest prime q with 11 <= q < 32768 dividing k*2^n + c, else 0.
 * Same result as try_all_factors (factor_proth.c:220-253), which computes
 * -c*2^-n mod q in 16-bit Montgomery form; here 2^n mod q is a plain
 * square-and-multiply. */
int orc_try_all_factors(uint64_t k, uint32_t n, int c) {
  pthread_once(&g_small_once, init_small_primes);
  for (uint32_t i = 0; i < g_num_small; i++) {
    uint32_t q = g_small_primes[i], t = 1, b = 2 % q, e = n;
    while (e) {
      if (e & 1) t = t * b % q;
      b = b * b % q;
      e >>= 1;
    }
    if (((uint64_t)(k % q) * t + (c > 0 ? 1u : q - 1u)) % q == 0) return (int)q;
  }
  return 0;
}

/* 1 where x = 0..14 is divisible by neither 3 nor 5 (app.c:150-152) */
static const int prime15[15] = {0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 1, 1};

/* Restates test_factor (app.c:159-186) in no-input-file mode: the k residue
 * rule, the n range, the shifted-k residue pre-tests with their k-size
 * guards (all in wrapping uint64 arithmetic as in C), then try_all_factors.
 * Returns 1 when the factor would be reported. */
int orc_test_factor(const orc_params *pr, uint64_t k, uint32_t n, int c) {
  uint64_t b = k / pr->kstep;
  const uint64_t cc = (uint64_t)(int64_t)c;
  uint32_t khigh;
  uint64_t mod31 = 1;
  int keep;
  if (!(k == (uint64_t)pr->kstep * b + pr->koffset && n >= pr->nmin && n < pr->nmax)) return 0;
  khigh = (uint32_t)(k >> 32);
  keep = prime15[(uint32_t)(((k << (n & 3)) + cc) % 15)] && (uint32_t)(((k << (n % 3)) + cc) % 7) != 0;
  if (keep && khigh < (1u << 24)) {
    keep = (uint32_t)(((k << (n & 7)) + cc) % 17) != 0;
    if (keep && khigh < (1u << 22)) {
      mod31 = (k << (n % 10)) + cc;
      keep = (uint32_t)(mod31 % 11) != 0;
      if (keep && khigh < (1u << 21)) {
        keep = (uint32_t)(((k << (n % 11)) + cc) % 23) != 0;
        if (keep && khigh < (1u << 20)) {
          keep = (uint32_t)(((k << (n % 12)) + cc) % 13) != 0;
          if (keep && khigh < (1u << 14)) keep = (uint32_t)(((k << (n % 18)) + cc) % 19) != 0;
        }
      }
    }
  }
  if (!keep) return 0;
  return (uint32_t)(mod31 % 31) != 0 && orc_try_all_factors(k, n, c) == 0;
}

/* ------------------------------------------------------------------ */
/* the walk                                                             */
/* ------------------------------------------------------------------ */

typedef struct {
  orc_hit *v;
  uint64_t n, cap;
} hitvec;

static void hv_push(hitvec *h, uint64_t p, uint64_t k, uint32_t n, int c) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 64;
    h->v = realloc(h->v, h->cap * sizeof(orc_hit));
  }
  h->v[h->n].p = p;
  h->v[h->n].k = k;
  h->v[h->n].n = n;
  h->v[h->n].c = c;
  h->n++;
}

/* One prime, one sign.  Restates test_one_p (app.c:1037-1121) over the whole
 * [nmin, nmax) range: K = -c * 2^-n mod p; in window [n, n+s) a solution
 * k*2^(n+i) + c == 0 (mod p) with k <= kmax means K == k*2^i exactly, found
 * as i = ctz(K), k = K >> i.  Raw hits keep k in [kmin,kmax] and n+i in
 * [nmin,nmax); with filter set, test_factor decides. */
static void walk_one(const orc_params *pr, uint64_t p, int c, int filter, hitvec *h) {
  const uint32_t s = pr->nstep;
  uint64_t ps = orc_neg_inv64(p);
  uint64_t K = orc_inv2pow(p, pr->nmin);
  uint32_t n = pr->nmin;
  if (c == 1) K = p - K;
  do {
    int i = __builtin_ctzll(K);
    uint64_t k = K >> i;
    if (k <= pr->kmax && k >= pr->kmin && (uint32_t)i < s) {
      uint32_t nn = n + (uint32_t)i;
      if (filter ? orc_test_factor(pr, k, nn, c) : (nn >= pr->nmin && nn < pr->nmax)) hv_push(h, p, k, nn, c);
    }
    n += s;
    K = orc_step(K, p, ps, s);
  } while (n < pr->nmax);
}

typedef struct {
  const orc_params *pr;
  const uint64_t *P;
  uint64_t lo, hi;
  int filter;
  hitvec h;
} walk_job;

static void *walk_thread(void *arg) {
  walk_job *j = arg;
  for (uint64_t i = j->lo; i < j->hi; i++) {
    if (j->pr->sign >= 0) walk_one(j->pr, j->P[i], 1, j->filter, &j->h);
    if (j->pr->sign <= 0) walk_one(j->pr, j->P[i], -1, j->filter, &j->h);
  }
  return NULL;
}

/* Walk every prime in P (any order) and return the hits in prime order.
 * Returns the total number of hits; at most cap are written. */
int64_t orc_walk(const orc_params *pr, const uint64_t *P, uint64_t np, orc_hit *out, uint64_t cap, int filter,
                 int nthreads) {
  walk_job *jobs;
  pthread_t *tid;
  uint64_t total = 0;
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > np) nthreads = np ? (int)np : 1;
  jobs = calloc((size_t)nthreads, sizeof(walk_job));
  tid = calloc((size_t)nthreads, sizeof(pthread_t));
  for (int t = 0; t < nthreads; t++) {
    jobs[t].pr = pr;
    jobs[t].P = P;
    jobs[t].lo = np * t / nthreads;
    jobs[t].hi = np * (t + 1) / nthreads;
    jobs[t].filter = filter;
    if (nthreads == 1)
      walk_thread(&jobs[t]);
    else
      pthread_create(&tid[t], NULL, walk_thread, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) {
    if (nthreads > 1) pthread_join(tid[t], NULL);
    for (uint64_t i = 0; i < jobs[t].h.n; i++, total++)
      if (total < cap) out[total] = jobs[t].h.v[i];
    free(jobs[t].h.v);
  }
  free(jobs);
  free(tid);
  return (int64_t)total;
}

/* Primes of [p0,p1) then the walk; count and sum of the primes as printed by
 * the reference at the end of a run (main.c:1224-1232). */
int64_t orc_run_range(const orc_params *pr, uint64_t p0, uint64_t p1, orc_hit *out, uint64_t cap, int filter,
                      int nthreads, uint64_t *count, uint64_t *sum) {
  int64_t np = orc_primes(p0, p1, NULL, 0), nh;
  uint64_t *P = malloc(sizeof(uint64_t) * (size_t)(np + 1)), s = 0;
  orc_primes(p0, p1, P, (uint64_t)np);
  for (int64_t i = 0; i < np; i++) s += P[i];
  if (count) *count = (uint64_t)np;
  if (sum) *sum = s;
  nh = orc_walk(pr, P, (uint64_t)np, out, cap, filter, nthreads);
  free(P);
  return nh;
}

/* "%llu | %llu*2^%u%+d\n" as report_factor writes it (app.c:134) */
/* Checks of reported divisors, independent of the walk: for each hit, bit
 * ORC_BAD_PRIME if p is not prime (Miller-Rabin), ORC_BAD_DIVIDES if
 * k*2^n + c is not 0 mod p (square-and-multiply in 128-bit arithmetic),
 * ORC_BAD_K if k is not odd in [kmin, kmax], ORC_BAD_N if n is outside
 * [nmin, nmax), ORC_BAD_SIGN if c is not the search's sign (+1 or -1 for
 * both), and ORC_KEPT if orc_test_factor keeps it.  Returns the number of
 * hits with any ORC_BAD_* bit. */
struct check_job {
  const orc_params *pr;
  const orc_hit *h;
  uint64_t lo, hi, bad;
  uint8_t *flags;
};

static void *check_thread(void *arg) {
  struct check_job *J = arg;
  const orc_params *pr = J->pr;
  for (uint64_t i = J->lo; i < J->hi; i++) {
    const orc_hit *x = &J->h[i];
    uint8_t f = 0;
    if (!orc_is_prime(x->p)) f |= ORC_BAD_PRIME;
    if (x->p > 2) {
      const uint64_t t = mulmod64(x->k % x->p, powmod64(2, x->n, x->p), x->p);
      const uint64_t want = x->c == 1 ? x->p - 1 : 1; /* k*2^n == -c (mod p) */
      if (t != want) f |= ORC_BAD_DIVIDES;
    } else
      f |= ORC_BAD_DIVIDES;
    if (!(x->k & 1) || x->k < pr->kmin || x->k > pr->kmax) f |= ORC_BAD_K;
    if (x->n < pr->nmin || x->n >= pr->nmax) f |= ORC_BAD_N;
    if (!((x->c == 1 || x->c == -1) && (pr->sign == 0 || x->c == pr->sign))) f |= ORC_BAD_SIGN;
    if (f) J->bad++;
    if (!f && orc_test_factor(pr, x->k, x->n, x->c)) f |= ORC_KEPT;
    J->flags[i] = f;
  }
  return NULL;
}

int64_t orc_check_hits(const orc_params *pr, const orc_hit *h, uint64_t n, int nthreads, uint8_t *flags) {
  struct check_job jobs[64];
  pthread_t tid[64];
  int64_t bad = 0;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  for (int t = 0; t < nthreads; t++) {
    jobs[t].pr = pr;
    jobs[t].h = h;
    jobs[t].lo = n * t / nthreads;
    jobs[t].hi = n * (t + 1) / nthreads;
    jobs[t].bad = 0;
    jobs[t].flags = flags;
    pthread_create(&tid[t], NULL, check_thread, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) {
    pthread_join(tid[t], NULL);
    bad += (int64_t)jobs[t].bad;
  }
  return bad;
}

int orc_format_hit(const orc_hit *h, char *buf, int len) {
  return snprintf(buf, (size_t)len, "%" PRIu64 " | %" PRIu64 "*2^%u%+d\n", h->p, h->k, h->n, h->c);
}
