/* ppsd.h -- C ABI of the MI355X (gfx950) PPSieve hot path: libppsd.so.
 *
 * This is the device layer that a PPSieve host calls where the reference's
 * app.c calls its CUDA layer through pps/appcu.h.  Plain C types only; one
 * context per device per host thread; no exit() inside the library: every
 * entry point returns 0 or a negative code, with the text from
 * ppsd_last_error().
 *
 * Semantics of a run (identical to the reference's no-input-file mode):
 * report every (p, k, n, c) with p prime in [pmin, pmax), k == koffset
 * (mod kstep) in [kmin, kmax], nmin <= n < nmax and p | k*2^n + c, where
 * k*2^n + c has no prime factor 3 <= q < 32768 (pps/app.c:150-186,
 * pps/factor_proth.c:246-253).  c = +1 (Proth), -1 (Riesel, -R) or both.
 */
#ifndef PPSD_H
#define PPSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes, BOINC-compatible like the reference's (pps/app.h:64-70). */
#define PPSD_OK 0
#define PPSD_ERR_FAILURE (-1)                /* EXIT_FAILURE paths: main.c, checkCUDAErr appcu.cu:124-131 */
#define PPSD_ERR_INVALID_PARAM (-178)        /* ERR_INVALID_PARAM, pps/app.h:67 */
#define PPSD_ERR_INSUFFICIENT_RESOURCE (-198) /* ERR_INSUFFICIENT_RESOURCE, pps/app.h:69 */
#define PPSD_ERR_NOT_IMPLEMENTED (-150)      /* ERR_NOT_IMPLEMENTED, pps/app.h:70 */
#define PPSD_ERR_STATE (-117)                /* ERR_NEG, pps/app.h:68: call out of order */

/* Search parameters.  Inputs use the reference's option meanings (main.c
 * -p/-P, pps/app.c -k/-K/-n/-N/-M/-s/-R); 0 means "not given".
 * ppsd_derive_params fills the outputs. */
typedef struct ppsd_params {
  uint64_t pmin, pmax;     /* -p, -P */
  uint64_t kmin, kmax;     /* -k, -K */
  uint64_t qmax;           /* -Q: sieve p with odd primes <= qmax only (0 = all <= sqrt) */
  uint32_t nmin, nmax;     /* -n, -N; nmax is exclusive */
  uint32_t kstep, koffset; /* -M, -s; kstep 0 selects 2 / 1 (odd k) */
  int32_t sign;            /* +1 Proth, -1 Riesel (-R), 0 both */
  uint32_t nstep_ref;      /* out: the reference's nstep after its clamps */
  uint32_t nstep;          /* out: window walked, min(nstep_ref, 32) */
  int32_t status;          /* out: 0 or PPSD_ERR_* */
  char msg[160];           /* out: reason for a non-zero status */
} ppsd_params;

/* One factor line "p | k*2^n+c" (pps/app.c:134). */
typedef struct ppsd_factor {
  uint64_t p;
  uint64_t k;
  uint32_t n;
  int32_t c;
} ppsd_factor;

/* Counters of the work since ppsd_open (or the last ppsd_reset_stats). */
typedef struct ppsd_stats {
  uint64_t primes;        /* primes walked */
  uint64_t prime_sum;     /* their sum mod 2^64 (main.c:1232 "sum=") */
  uint64_t window_steps;  /* primes x windows x signs walked */
  uint64_t raw_hits;      /* divisors found before the small-prime filter */
  uint64_t factors;       /* factors returned */
  uint64_t walk_launches;
  double walk_ms;         /* walk kernels, HIP events on the context's stream */
  double sieve_ms;        /* on-device prime supply kernels */
  double filter_ms;       /* small-prime filter kernel */
  uint64_t retries;       /* batches re-walked in halves after a full hit ring */
} ppsd_stats;

typedef struct ppsd_ctx ppsd_ctx;

/* Parameter rules: pmin/pmax defaults (main.c:916-932), validation, nmin
 * default, k alignment and nstep (pps/app.c:655-712), nstep clamps and range
 * checks (pps/appcu.cu:155-209, 236-239).  Replaces the checks inside
 * app_init + cuda_init (pps/appcu.h:31).  Returns p->status. */
int ppsd_derive_params(ppsd_params *p);

/* Number of HIP devices (0 without a GPU). */
int ppsd_device_count(void);

/* Peak shader clock of a device in MHz (its reported clock rate). */
int ppsd_device_clock(int device, uint32_t *max_mhz);

/* Architecture name (e.g. "gfx950:sramecc+:xnack-") and compute-unit count
 * of a device; name may be NULL. */
int ppsd_device_info(int device, char *name, uint32_t name_len, uint32_t *cus);

/* Per-worker device set-up; replaces cuda_app_init (pps/appcu.h:27,
 * pps/appcu.cu:213-404).  p normally comes from ppsd_derive_params; any p
 * with status 0, 2^32 < pmin < pmax <= 2^62, 64 <= nmin <= nmax,
 * 1 <= kmin <= kmax, 1 <= nstep <= 32 and kmax * 2^(nstep-1) < pmin (one
 * candidate per window) is accepted, which also admits work units outside
 * the reference's own validation (2^(64-nstep) <= pmin).  batch is
 * the number of primes per walk launch (0 = default: 32 per SIMD lane slot of
 * the device, i.e. 2^21 on 256 CUs).  Allocates every device buffer up
 * front; generates the sieving primes <= isqrt(pmax-1) on the device. */
int ppsd_open(int device, const ppsd_params *p, uint32_t batch, ppsd_ctx **out);

/* Primes per batch chosen by ppsd_open (the reference's cthread_count). */
uint32_t ppsd_batch_size(const ppsd_ctx *ctx);

/* Walk the n-range for n host-supplied primes (n <= batch; any order),
 * asynchronously; replaces check_ns (pps/appcu.h:29, pps/appcu.cu:1205-1261).
 * P is copied before return.  One submission may be outstanding. */
int ppsd_submit_primes(ppsd_ctx *ctx, const uint64_t *P, uint32_t n);

/* On-device prime supply: sieve [p0, p1) on the device and walk its primes,
 * asynchronously.  At most one batch of primes is taken: *p_next receives
 * where this submission stopped (p1 when the whole range fitted).
 * No counterpart in the reference, whose primes come from the host sieve
 * (sieve.c, main.c:750-777). */
int ppsd_submit_range(ppsd_ctx *ctx, uint64_t p0, uint64_t p1, uint64_t *p_next);

/* Wait for the outstanding submission and return its factors (filtered as
 * test_factor does); replaces get_factors_found (pps/appcu.h:30,
 * pps/appcu.cu:1263-1308) plus the host re-walk it fed (pps/app.c:1197-1219).
 * Writes at most cap factors; *n_out receives the total (call again with a
 * larger buffer when *n_out > cap: the result is kept until the next submit).
 * primes/prime_sum (may be NULL) receive the batch's count and sum. */
int ppsd_collect(ppsd_ctx *ctx, ppsd_factor *out, uint64_t cap, uint64_t *n_out, uint64_t *primes,
                 uint64_t *prime_sum);

/* Same wait, returning the raw divisors instead (no small-prime filter):
 * every (p,k,n,c) with p | k*2^n+c, k odd in [kmin,kmax], nmin <= n < nmax.
 * Either this or ppsd_collect consumes a submission. */
int ppsd_collect_raw(ppsd_ctx *ctx, ppsd_factor *out, uint64_t cap, uint64_t *n_out, uint64_t *primes,
                     uint64_t *prime_sum);

/* Kernel shape: r primes per lane (1, 2), u windows per pre-test chunk
 * (4, 8, 16), workgroup size (64..256, multiple of 64).  0 keeps a value. */
int ppsd_set_tuning(ppsd_ctx *ctx, int r, int u, int wg);
int ppsd_get_tuning(const ppsd_ctx *ctx, int *r, int *u, int *wg);

/* Form of the walk's per-window pre-test (no counterpart in the reference,
 * whose kernels pick among test macros at compile time: pps/appcu.cu:779-801).
 * 1: the odd part of the state is <= kmax (ctz, shift, compare); 2: the
 * state's high word is <= that of kmax*2^(s-1) (or, both signs, >= that of
 * p minus it), a running minimum (maximum) over a chunk of windows -- a
 * chunk re-walk whenever a merely small state fires it; 3:
 * the low 32 bits of the odd part, minimised over a chunk of windows, are <=
 * kmax -- fires on about kmax/2^32 of all states, for kmax < 2^32 only (else
 * 1 is used); 0 (default): the form, and a chunk
 * no longer than the tuned u, that a cost model of their re-walks prices
 * lowest.  Results never depend on it.  plain_form and plain_u receive the
 * form (1, 2 or 3) and the chunk length this context's own walk uses. */
int ppsd_set_pretest(ppsd_ctx *ctx, int mode);
int ppsd_get_pretest(const ppsd_ctx *ctx, int *mode, int *plain_form, int *plain_u);

/* The form (1, 2 or 3, as above) and chunk length the automatic choice
 * gives a unit (parameters as ppsd_open takes them) at kernel shape (r, u):
 * for the plain walk (batched 0), which may shorten the chunk below u, or
 * the batched walk (batched 1).  No device is needed. */
int ppsd_walk_form(const ppsd_params *p, int r, int u, int batched, int *form, int *chunk);

/* Raw-hit ring capacity for later submissions (1 .. the allocated 2^24).
 * A batch whose raw hits overflow the ring is re-walked in halves, so the
 * capacity only trades memory for re-walks; the result never changes. */
int ppsd_set_hit_capacity(ppsd_ctx *ctx, uint64_t cap);

/* Candidate list of a sieve file (-i; pps/app.c:181-184, read by
 * pps/app.c:349-517): when set, ppsd_collect keeps the divisors whose
 * (k, n) is a listed candidate instead of those passing the small-prime
 * rule.  k = kstep*b + koffset is listed when bit (b - b0) % 8 of byte
 * (b - b0) / 8 in row (n - n0) is set; rows rows of row_bytes bytes.  The
 * library keeps a device copy; bitmap NULL clears it. */
int ppsd_set_candidates(ppsd_ctx *ctx, const uint8_t *bitmap, uint32_t n0, uint32_t rows, uint64_t b0,
                        uint64_t row_bytes);

int ppsd_get_stats(const ppsd_ctx *ctx, ppsd_stats *st);
int ppsd_reset_stats(ppsd_ctx *ctx);

/* Frees all device state; replaces cuda_finalize (pps/appcu.h:32,
 * pps/appcu.cu:101-122).  NULL is accepted. */
int ppsd_close(ppsd_ctx *ctx);

/* Text of the last error on this thread. */
const char *ppsd_last_error(void);

/* ---- Batched search (no counterpart in the reference, whose device layer
 * runs one search per process: pps/appcu.cu:1205-1308).  For callers with
 * many independent searches -- sieving farms handing out small work units,
 * the serving loop -- that would each leave most of the GPU idle.
 *
 * ppsd_batch_run searches every unit completely: each is a full ppsd_params
 * (its own p-, k- and n-range, sign, window, k filter) accepted by ppsd_open,
 * with no partial sieve (qmax 0 or >= isqrt(pmax-1)).  The units' primes are
 * sieved on the device a round at a time into prime pools (the next rounds
 * beside the current one's walk); one persistent kernel then drains a device
 * queue of (unit, 64*r primes, n-segment) tasks, so that the whole GPU works
 * on whatever mix of units is present.  Results are those of
 * searching each unit alone with ppsd_submit_range: the same factors (or,
 * raw != 0, the same divisors before test_factor) -- the order differs.
 * Synchronous; the context must have no outstanding submission and must not
 * have been opened with a partial sieve.  Counters go to ppsd_get_stats. */
int ppsd_batch_run(ppsd_ctx *ctx, const ppsd_params *units, uint32_t nunits, int raw);

/* The batched path's one-time set-up, done ahead of the first ppsd_batch_run
 * that would otherwise do it: its device buffers (about 6.4 GB) and pinned
 * host buffers (about 560 MB, most of it the mirror the results are copied
 * to during the run), the sieving primes for every unit with pmax <= pmax_all, and the
 * round graphs of the current kernel shape and ring capacity.  Optional:
 * ppsd_batch_run does whatever is missing.  No counterpart in the reference,
 * whose host allocates per device in cuda_app_init (pps/appcu.cu:213-404). */
int ppsd_batch_prepare(ppsd_ctx *ctx, uint64_t pmax_all);

/* Hits of the last ppsd_batch_run: up to cap of them to out, with the index
 * of their unit in unit_of (either may be NULL); *n_out = their number. */
int ppsd_batch_results(ppsd_ctx *ctx, ppsd_factor *out, uint32_t *unit_of, uint64_t cap, uint64_t *n_out);

/* Per unit of the last ppsd_batch_run: primes walked and their sum mod 2^64
 * (the run's checksum), for the first nunits units. */
int ppsd_batch_unit_primes(const ppsd_ctx *ctx, uint64_t *count, uint64_t *sum, uint32_t nunits);

/* Shape of the batched kernel: r primes per lane (1, 2), u windows per
 * pre-test chunk (4, 8, 16), resident waves per SIMD (1..16), and the
 * n-segment length in windows, chosen per run between seg_min and seg_max
 * (seg_min >= 16).  0 keeps a value.  Defaults: r = 1, u = 16, 6 waves per
 * SIMD, segments of 2048..65536 windows. */
int ppsd_batch_set_tuning(ppsd_ctx *ctx, int r, int u, int waves_per_simd, uint32_t seg_min, uint32_t seg_max);
int ppsd_batch_get_tuning(const ppsd_ctx *ctx, int *r, int *u, int *waves_per_simd, uint32_t *seg_min,
                          uint32_t *seg_max);

/* Raw-hit ring capacity of the batched walk (1 .. 2^25); overflowing task
 * sets are re-run in halves, so results never depend on it.  The batch's
 * device output then holds at most twice this many results before they are
 * moved to host memory (results do not depend on that either). */
int ppsd_batch_set_hit_capacity(ppsd_ctx *ctx, uint64_t cap);

/* ---- Serving loop (no counterpart in the reference).  For callers that
 * receive searches continuously (a work-unit server) and need each answered
 * soon after it arrives.  One server owns one device and a worker thread.
 *
 * Requests are full ppsd_params, as for ppsd_batch_run.  They queue in
 * arrival order and are admitted while the device memory they hold -- their
 * share of the resident prime pool (primes of their next p-piece, at most a
 * quarter of the pool) plus a hit-ring reservation -- fits mem_budget.
 * The device then runs slices of about slice_ms (twice that while requests
 * wait for admission): each slice is one round of
 * the batched walk over one p-piece of every admitted request, the slice's
 * work shared equally among them, smallest remaining work first.  A request
 * arriving during a slice joins the next one.  Each request's result equals
 * searching it alone (ppsd_submit_range): same factors (or raw divisors),
 * same prime count and sum.  Times are seconds since ppsd_serve_open. */
typedef struct ppsd_server ppsd_server;

typedef struct ppsd_serve_config {
  uint64_t mem_budget; /* bytes admitted requests may hold (0 = pool + rings, about 0.8 GiB) */
  double slice_ms;     /* target length of one slice (0 = 50); twice that while requests wait */
  uint32_t max_active; /* admitted requests at most (0 = 4096; <= 65536) */
  int32_t raw;         /* 1: raw divisors instead of filtered factors */
} ppsd_serve_config;

typedef struct ppsd_serve_result {
  uint64_t id;
  int32_t status; /* 0, or the error code that ended the request */
  uint32_t slices;
  uint64_t primes, prime_sum, factors, window_steps;
  double t_arrival, t_admit, t_start, t_done; /* queue = admit - arrival .. start; service = done - start */
} ppsd_serve_result;

typedef struct ppsd_serve_stats {
  uint64_t submitted, admitted, completed, queued, active;
  uint64_t slices, primes, window_steps;
  uint64_t mem_in_use, mem_peak, mem_budget;
  double busy_s; /* time spent in slices */
  double now;
} ppsd_serve_stats;

/* Opens device `device` (cfg may be NULL) and starts the worker. */
int ppsd_serve_open(int device, const ppsd_serve_config *cfg, ppsd_server **out);
/* Queues a request (thread-safe); *id identifies its result. */
int ppsd_serve_submit(ppsd_server *s, const ppsd_params *req, uint64_t *id);
/* Takes the next completed request, waiting up to timeout_ms: 0 = a result
 * in *res, 1 = none within the timeout. */
int ppsd_serve_next(ppsd_server *s, int timeout_ms, ppsd_serve_result *res);
/* Factors of a taken result: up to cap to out, *n_out = their number. */
int ppsd_serve_factors(ppsd_server *s, uint64_t id, ppsd_factor *out, uint64_t cap, uint64_t *n_out);
/* Frees a taken result. */
int ppsd_serve_release(ppsd_server *s, uint64_t id);
int ppsd_serve_get_stats(ppsd_server *s, ppsd_serve_stats *st);
/* Stops the worker after its current slice; unfinished requests are dropped. */
int ppsd_serve_close(ppsd_server *s);

/* ---- Allocation-free walk.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It is the
 * one-call, context-free form of ppsd_submit_primes followed by
 * ppsd_collect (raw == 0) or ppsd_collect_raw (raw != 0), on primes and an
 * output buffer that are already in device memory, and it allocates
 * nothing: its raw-hit ring (PPSD_NOALLOC_RING_HITS records), counters and
 * the filter's small-prime table with its powers of two live in a fixed
 * device array of the library (about 15.4 MB per GPU, held until the
 * process ends).
 *
 * P: n device-resident primes, odd in [pmin, 2^62), any order, n <=
 * PPSD_NOALLOC_MAX_PRIMES (more is PPSD_ERR_INVALID_PARAM).  out: a device
 * buffer of cap ppsd_factor records; the first min(*n_out, cap) results are
 * written there in no particular order, *n_out receives their number.
 * stream: a hipStream_t to run on (NULL: the device's null stream); the call
 * returns when the work is done.  p is used as ppsd_open takes it, with the
 * library's default kernel shape.  Every call on one GPU uses that GPU's one
 * array: calls that overlap on one GPU race and must be serialised by the
 * caller.  A prime with more than PPSD_NOALLOC_RING_HITS raw divisors is
 * PPSD_ERR_INSUFFICIENT_RESOURCE. */
#define PPSD_NOALLOC_RING_HITS (1u << 19)
#define PPSD_NOALLOC_MAX_PRIMES (1u << 24)
int ppsd_walk_primes_noalloc(int device, const ppsd_params *p, const uint64_t *P, uint32_t n, ppsd_factor *out,
                             uint64_t cap, int raw, void *stream, uint64_t *n_out);

/* Arithmetic check (no counterpart in the reference): out[i] = 2^-e mod P[i]
 * after nwin further windows of s, i.e. 2^-(e + nwin*s) mod P[i], computed by
 * the same device code as the walk, returned as the walk holds it (the
 * walk keeps every state canonical, in [0, P[i])).  Requires P[i] odd in
 * (2^32, 2^62), e >= 64, 1 <= s <= 32. */
int ppsd_debug_state(int device, const uint64_t *P, uint32_t n, uint32_t e, uint32_t nwin, uint32_t s,
                     uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* PPSD_H */
