/* pps_hostsieve.c -- host segmented sieve for `ppsieve --host-primes`.
 *
 * The candidates are the odd numbers of [p0, p1): index j is the value
 * lo + 2j (lo = p0 | 1), the odd-only mapping of the reference's bitmap
 * (main.c:766).  Segments of 2^21 candidates (256 KiB of bits, one
 * composite flag each) are sieved in turn:
 *   - 3, 5, 7, 11 and 13 by copying a precomputed pattern of period
 *     3*5*7*11*13 = 15015 candidates at the segment's phase;
 *   - primes below the segment length by stepping from a per-prime next
 *     index carried from segment to segment;
 *   - larger primes, which hit a segment at most once, from buckets: each
 *     sits in the bucket of the segment holding its next multiple and moves
 *     on after marking it, so no prime is visited by a segment it misses.
 *     The buckets form a ring over the next few segments (one stride of the
 *     largest prime); a first multiple further ahead (q^2 far above the
 *     range start) waits in a list sorted by index until the ring reaches it.
 * Only setting up a range takes a remainder per sieving prime, by a
 * reciprocal computed once per prime.
 *
 * With several threads, a range is cut into as many contiguous parts: the
 * calling thread sieves the first and hands its primes on segment by
 * segment, while helper threads sieve the others into buffers that are
 * handed on in order afterwards.
 */
#include "pps_hostsieve.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define SEG_LOG 21
#define SEG ((uint64_t)1 << SEG_LOG)
#define SEG_WORDS (SEG / 64)
#define PAT_PERIOD 15015u /* 3*5*7*11*13 */
#define PAT_WORDS ((PAT_PERIOD + 128) / 64 + 1)
#define MAX_PARTS 64

typedef struct {
  uint64_t *e; /* entries: prime index << SEG_LOG | offset in the segment */
  size_t n, cap;
} bucket;

typedef struct {
  uint64_t j; /* candidate index of the first multiple */
  uint32_t i; /* prime index */
} later;

/* The sieving state of one thread. */
typedef struct {
  uint64_t *bits; /* one segment */
  uint64_t *out;  /* primes of one segment */
  uint64_t *next; /* medium primes: next multiple's index */
  bucket *bk;
  later *lt; /* first multiples beyond the ring, by index */
  size_t nlt, caplt;
} part_state;

struct pps_hostsieve {
  const uint32_t *q;
  uint64_t *inv; /* floor((2^64 - 1) / q) */
  uint32_t nq;
  uint64_t pat[PAT_WORDS]; /* bit i: 1 + 2i has a factor among 3..13 */
  uint32_t nbk;            /* ring buckets: power of two above the widest stride in segments */
  int nthreads;
  part_state *ps;
};

/* A helper thread's part: its primes, gathered for handing on in order. */
typedef struct {
  pps_hostsieve *h;
  part_state *st;
  uint64_t p0, p1;
  uint64_t *buf;
  size_t n, cap;
  int rc;
} part_job;

static int push(bucket *b, uint64_t v) {
  if (b->n == b->cap) {
    size_t cap = b->cap ? 2 * b->cap : 256;
    uint64_t *e = realloc(b->e, cap * sizeof *e);
    if (!e) return -1;
    b->e = e;
    b->cap = cap;
  }
  b->e[b->n++] = v;
  return 0;
}

static void part_free(part_state *st, uint32_t nbk) {
  if (st->bk)
    for (uint32_t i = 0; i < nbk; i++) free(st->bk[i].e);
  free(st->bk);
  free(st->lt);
  free(st->bits);
  free(st->out);
  free(st->next);
}

pps_hostsieve *pps_hostsieve_new(const uint32_t *q, uint32_t nq, int nthreads) {
  pps_hostsieve *h = calloc(1, sizeof *h);
  if (!h) return NULL;
  h->q = q;
  h->nq = nq;
  h->nthreads = nthreads < 1 ? 1 : nthreads;
  for (uint32_t i = 0; i < PAT_PERIOD + 128; i++) {
    const uint32_t v = 1 + 2 * (i % PAT_PERIOD);
    if (v % 3 == 0 || v % 5 == 0 || v % 7 == 0 || v % 11 == 0 || v % 13 == 0) h->pat[i / 64] |= (uint64_t)1 << (i % 64);
  }
  const uint64_t qmax = nq ? q[nq - 1] : 1;
  h->nbk = 2;
  while ((uint64_t)h->nbk < qmax / SEG + 2) h->nbk *= 2;
  h->inv = malloc((size_t)(nq ? nq : 1) * sizeof *h->inv);
  h->ps = calloc((size_t)h->nthreads, sizeof *h->ps);
  if (!h->inv || !h->ps) {
    pps_hostsieve_free(h);
    return NULL;
  }
  for (uint32_t i = 0; i < nq; i++) h->inv[i] = UINT64_MAX / q[i];
  for (int t = 0; t < h->nthreads; t++) {
    part_state *st = &h->ps[t];
    st->bits = malloc(SEG_WORDS * sizeof *st->bits);
    st->out = malloc(SEG * sizeof *st->out);
    st->next = malloc((size_t)(nq ? nq : 1) * sizeof *st->next);
    st->bk = calloc(h->nbk, sizeof *st->bk);
    if (!st->bits || !st->out || !st->next || !st->bk) {
      pps_hostsieve_free(h);
      return NULL;
    }
  }
  return h;
}

void pps_hostsieve_free(pps_hostsieve *h) {
  if (!h) return;
  if (h->ps)
    for (int t = 0; t < h->nthreads; t++) part_free(&h->ps[t], h->nbk);
  free(h->ps);
  free(h->inv);
  free(h);
}

/* x mod q from inv = floor((2^64 - 1) / q): the estimate is the quotient
 * or one less, so one conditional subtraction finishes. */
static inline uint64_t mod_inv(uint64_t x, uint64_t q, uint64_t inv) {
  uint64_t r = x - (uint64_t)(((unsigned __int128)x * inv) >> 64) * q;
  return r >= q ? r - q : r;
}

static int by_index(const void *a, const void *b) {
  const uint64_t x = ((const later *)a)->j, y = ((const later *)b)->j;
  return x < y ? -1 : x > y;
}

/* 64 pattern bits from bit o (o < PAT_PERIOD) */
static inline uint64_t pat_word(const uint64_t *pat, uint32_t o) {
  const uint32_t w = o / 64, b = o % 64;
  return b ? (pat[w] >> b) | (pat[w + 1] << (64 - b)) : pat[w];
}

/* One contiguous part [p0, p1) on one thread's state. */
static int sieve_part(const pps_hostsieve *h, part_state *st, uint64_t p0, uint64_t p1,
                      void (*emit)(void *arg, const uint64_t *primes, size_t n), void *arg) {
  const uint64_t lo = p0 | 1;
  if (lo >= p1) return 0;
  const uint64_t ncand = (p1 - lo + 1) / 2;
  const uint64_t nseg = (ncand + SEG - 1) / SEG;
  const uint32_t mask = h->nbk - 1;
  for (uint32_t i = 0; i < h->nbk; i++) st->bk[i].n = 0;
  st->nlt = 0;
  /* primes from 17 with q^2 < p1: first odd multiple >= max(q^2, lo) */
  uint32_t i0 = 0, nmed = 0, nq = 0;
  while (i0 < h->nq && h->q[i0] < 17) i0++;
  for (nq = i0; nq < h->nq && (uint64_t)h->q[nq] * h->q[nq] < p1; nq++) {
    const uint64_t q = h->q[nq], qq = q * q;
    uint64_t m = qq;
    if (qq < lo) {
      const uint64_t r = mod_inv(lo, q, h->inv[nq]);
      m = r ? lo + (q - r) : lo;
      if (!(m & 1)) m += q;
    }
    const uint64_t j = (m - lo) / 2;
    if (q < SEG) {
      st->next[nq] = j;
      nmed = nq + 1;
    } else if (j < ncand && (j >> SEG_LOG) < h->nbk) {
      if (push(&st->bk[(j >> SEG_LOG) & mask], (uint64_t)nq << SEG_LOG | (j & (SEG - 1)))) return -1;
    } else if (j < ncand) {
      if (st->nlt == st->caplt) {
        size_t cap = st->caplt ? 2 * st->caplt : 1024;
        later *lt = realloc(st->lt, cap * sizeof *lt);
        if (!lt) return -1;
        st->lt = lt;
        st->caplt = cap;
      }
      st->lt[st->nlt++] = (later){j, nq};
    }
  }
  qsort(st->lt, st->nlt, sizeof *st->lt, by_index);
  size_t lt_next = 0;
  for (uint64_t t = 0; t < nseg; t++) {
    /* first multiples the ring now reaches: segments below t + nbk */
    for (; lt_next < st->nlt && (st->lt[lt_next].j >> SEG_LOG) < t + h->nbk; lt_next++) {
      const uint64_t j = st->lt[lt_next].j;
      if (push(&st->bk[(j >> SEG_LOG) & mask], (uint64_t)st->lt[lt_next].i << SEG_LOG | (j & (SEG - 1)))) return -1;
    }
    const uint64_t s0 = t * SEG;
    const uint64_t len = ncand - s0 < SEG ? ncand - s0 : SEG;
    const uint64_t v0 = lo + 2 * s0;
    const uint64_t nw = (len + 63) / 64;
    /* 3..13 */
    uint32_t o = (uint32_t)(((v0 - 1) / 2) % PAT_PERIOD);
    for (uint64_t w = 0; w < nw; w++) {
      st->bits[w] = pat_word(h->pat, o);
      o += 64;
      if (o >= PAT_PERIOD) o -= PAT_PERIOD;
    }
    if (v0 <= 13) /* the small primes themselves */
      for (uint64_t v = 3; v <= 13; v += 2)
        if (v >= v0 && (v == 3 || v == 5 || v == 7 || v == 11 || v == 13)) {
          const uint64_t b = (v - v0) / 2;
          st->bits[b / 64] &= ~((uint64_t)1 << (b % 64));
        }
    /* medium primes */
    const uint64_t s1 = s0 + len;
    for (uint32_t i = i0; i < nmed; i++) {
      const uint64_t q = h->q[i];
      uint64_t j = st->next[i];
      for (; j < s1; j += q) {
        const uint64_t b = j - s0;
        st->bits[b / 64] |= (uint64_t)1 << (b % 64);
      }
      st->next[i] = j;
    }
    /* large primes due in this segment */
    bucket *bk = &st->bk[t & mask];
    for (size_t e = 0; e < bk->n; e++) {
      const uint64_t ent = bk->e[e];
      const uint32_t i = (uint32_t)(ent >> SEG_LOG);
      const uint64_t b = ent & (SEG - 1);
      st->bits[b / 64] |= (uint64_t)1 << (b % 64);
      const uint64_t j = s0 + b + h->q[i];
      if (j < ncand && push(&st->bk[(j >> SEG_LOG) & mask], (uint64_t)i << SEG_LOG | (j & (SEG - 1)))) return -1;
    }
    bk->n = 0;
    /* the primes: zero bits below len */
    size_t n = 0;
    for (uint64_t w = 0; w < nw; w++) {
      uint64_t z = ~st->bits[w];
      if (w == nw - 1 && (len % 64)) z &= ((uint64_t)1 << (len % 64)) - 1;
      while (z) {
        const int b = __builtin_ctzll(z);
        z &= z - 1;
        st->out[n++] = v0 + 2 * (64 * w + (uint64_t)b);
      }
    }
    if (n) emit(arg, st->out, n);
  }
  return 0;
}

static void gather(void *arg, const uint64_t *p, size_t n) {
  part_job *j = arg;
  if (j->rc) return;
  if (j->n + n > j->cap) {
    size_t cap = j->cap ? j->cap : 1 << 20;
    while (cap < j->n + n) cap *= 2;
    uint64_t *b = realloc(j->buf, cap * sizeof *b);
    if (!b) {
      j->rc = -1;
      return;
    }
    j->buf = b;
    j->cap = cap;
  }
  memcpy(j->buf + j->n, p, n * sizeof *p);
  j->n += n;
}

static void *part_thread(void *arg) {
  part_job *j = arg;
  if (sieve_part(j->h, j->st, j->p0, j->p1, gather, j)) j->rc = -1;
  return NULL;
}

int pps_hostsieve_range(pps_hostsieve *h, uint64_t p0, uint64_t p1,
                        void (*emit)(void *arg, const uint64_t *primes, size_t n), void *arg) {
  if (p0 < 3) p0 = 3;
  if (p0 >= p1) return 0;
  /* parts of at least 2^26 numbers, so that each part's set-up (a remainder
   * per sieving prime) stays small beside its sieving */
  int np = h->nthreads < MAX_PARTS ? h->nthreads : MAX_PARTS;
  while (np > 1 && (p1 - p0) / (uint64_t)np < ((uint64_t)1 << 26)) np--;
  if (np == 1) return sieve_part(h, &h->ps[0], p0, p1, emit, arg);
  const uint64_t width = ((p1 - p0) / (uint64_t)np) & ~(uint64_t)1;
  part_job jobs[MAX_PARTS];
  pthread_t tid[MAX_PARTS];
  int started[MAX_PARTS] = {0};
  for (int k = 1; k < np; k++) {
    const uint64_t a = p0 + (uint64_t)k * width, b = k == np - 1 ? p1 : a + width;
    jobs[k] = (part_job){h, &h->ps[k], a, b, NULL, 0, 0, 0};
    started[k] = pthread_create(&tid[k], NULL, part_thread, &jobs[k]) == 0;
  }
  int rc = sieve_part(h, &h->ps[0], p0, p0 + width, emit, arg);
  for (int k = 1; k < np; k++) {
    if (started[k])
      pthread_join(tid[k], NULL);
    else
      part_thread(&jobs[k]); /* no thread could be started: sieve it here, in order */
    if (jobs[k].rc) rc = -1;
    for (size_t i = 0; !rc && i < jobs[k].n; i += SEG) emit(arg, jobs[k].buf + i, jobs[k].n - i < SEG ? jobs[k].n - i : SEG);
    free(jobs[k].buf);
  }
  return rc;
}
