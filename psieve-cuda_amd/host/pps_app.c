/* pps_app.c -- the Proth/Riesel application behind the PPSieve hooks,
 * running its hot path on an MI355X through libppsd (include/ppsd.h).
 *
 * Hook contract: pps/app.h:73-84 as main.c calls them.  What changes from
 * the reference's pps/app.c is where the work happens: the device returns
 * exact, already filtered factors (no per-prime flag byte and no host
 * re-walk, pps/app.c:1197-1219), and app_thread_range lets the device
 * generate the primes itself.
 */
#include <getopt.h>
#include <inttypes.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/ppsd.h"
#include "pps_app.h"
#include "pps_ckpt.h"
#include "pps_input.h"

uint64_t kmin = 0, kmax = 0;
unsigned int nmin = 0, nmax = 0;
#ifdef PPS_TWIN
int search_proth = 0; /* both signs, always */
#else
int search_proth = 1; /* +1 Proth, -1 Riesel (-R), 0 both (--both) */
#endif

static const char *input_filename = NULL;
static const char *factors_filename = NULL;
static FILE *factors_file = NULL;
static unsigned int factor_count = 0;
static int print_factors = 1;
static int device_opt = -1;
static unsigned int user_cthread_count = 0;
static unsigned int kstep = KSTEP_DEFAULT, koffset = KOFFSET_DEFAULT;
static ppsd_params params;
/* candidate list of the -i file: rows n = nmin..nmax, bit (k - kmin)/kstep */
static uint8_t *cand_bitmap = NULL;
static uint64_t cand_row_bytes = 0;
static uint32_t cand_rows = 0;
static pthread_mutex_t factors_mutex = PTHREAD_MUTEX_INITIALIZER;

#define MAX_WORKERS 32
/* Two contexts per worker, used in turn: a batch is checked, staged and
 * launched on one while the other's batch is still walking, so the device
 * does not wait for the host's share of a submission or of a collection.
 * ctxs[th][1] is NULL when a second context could not be opened. */
static ppsd_ctx *ctxs[MAX_WORKERS][2];
static int pending[MAX_WORKERS][2]; /* a submission is on the device */
static int next_ctx[MAX_WORKERS];   /* the context of the next submission */
static ppsd_factor *fbuf[MAX_WORKERS];
static uint64_t fcap[MAX_WORKERS];

/* One line per factor, "%llu | %llu*2^%u%+d" (pps/app.c:124-148). */
static void report_factor(const ppsd_factor *f) {
  pthread_mutex_lock(&factors_mutex);
  if (factors_file != NULL &&
      fprintf(factors_file, "%" PRIu64 " | %" PRIu64 "*2^%u%+d\n", f->p, f->k, f->n, f->c) > 0) {
    if (print_factors) printf("%" PRIu64 " | %" PRIu64 "*2^%u%+d\n", f->p, f->k, f->n, f->c);
  } else
    fprintf(stderr, "%sUNSAVED: %" PRIu64 " | %" PRIu64 "*2^%u%+d\n", bmprefix(), f->p, f->k, f->n, f->c);
  factor_count++;
  pthread_mutex_unlock(&factors_mutex);
}

static void device_fail(int th, const char *where, int rc) {
  fprintf(stderr, "%sDevice error (thread %d): %s: %s\n", bmprefix(), th, where, ppsd_last_error());
  bexit(rc == PPSD_ERR_INSUFFICIENT_RESOURCE ? ERR_INSUFFICIENT_RESOURCE : EXIT_FAILURE);
}

/* Wait for the submission on thread th's context j and report its factors. */
static int drain(int th, int j, uint64_t *count, uint64_t *sum) {
  uint64_t n = 0, cnt = 0, s = 0;
  int rc;
  if (!pending[th][j]) return 0;
  pending[th][j] = 0;
  rc = ppsd_collect(ctxs[th][j], fbuf[th], fcap[th], &n, &cnt, &s);
  if (rc) return rc;
  if (n > fcap[th]) {
    free(fbuf[th]);
    fcap[th] = n + n / 2;
    fbuf[th] = malloc(fcap[th] * sizeof(ppsd_factor));
    if (!fbuf[th]) return PPSD_ERR_INSUFFICIENT_RESOURCE;
    rc = ppsd_collect(ctxs[th][j], fbuf[th], fcap[th], &n, &cnt, &s);
    if (rc) return rc;
  }
  for (uint64_t i = 0; i < n; i++) report_factor(&fbuf[th][i]);
  if (count) *count += cnt;
  if (sum) *sum += s;
  return 0;
}

void app_banner(void) { printf(APP_NAME " version " APP_VERSION " (MI355X, gfx950)\n"); }

/* Option meanings of pps/app.c:551-617; returns 0, -1 invalid, -2 range. */
int app_parse_option(int opt, char *arg, const char *source) {
  int status = 0;
  switch (opt) {
    case 'k': status = parse_uint64(&kmin, arg, 1, (UINT64_C(1) << 62) - 1); break;
    case 'K': status = parse_uint64(&kmax, arg, 1, (UINT64_C(1) << 62) - 1); break;
    case 'm': status = parse_uint(&user_cthread_count, arg, 1, 1U << 31); break;
    case 'n': status = parse_uint(&nmin, arg, 1, (1U << 31) - 1); break;
    case 'N': status = parse_uint(&nmax, arg, 1, (1U << 31) - 1); break;
    case 'i': input_filename = (source == NULL) ? arg : strdup(arg); break;
    case 'f': factors_filename = (source == NULL) ? arg : strdup(arg); break;
    case 'd': status = parse_uint((unsigned int *)&device_opt, arg, 0, INT32_MAX); break;
    case 'M':
      status = parse_uint(&kstep, arg, 1, (1U << 31) - 1);
      if (koffset >= kstep) koffset = kstep / 2;
      break;
    case 's': status = parse_uint(&koffset, arg, 1, (1U << 31) - 1); break;
#ifndef PPS_TWIN
    case 'R': search_proth = -1; break;
    case 257: search_proth = 0; break;
#endif
    case 'b': break; /* accepted and ignored, as pps/app.c:557-561 */
  }
  return status;
}

void app_help(void) {
  printf("-f --factors=FILE  Write factors to FILE (default `%s')\n", FACTORS_FILENAME_DEFAULT);
  printf("-i --input=FILE    Read initial sieve from FILE\n");
  printf("-M --modulus=M     Print only k's == s mod M (default %u)\n", KSTEP_DEFAULT);
  printf("-s --modshift=S    (default %u, or M/2)\n", KOFFSET_DEFAULT);
  printf("-k --kmin=K0\n");
  printf("-K --kmax=K1       Sieve for primes k*2^n+/-1 with K0 <= k <= K1\n");
  printf("-m --mthreads=M    Primes per device batch = M per compute unit\n");
  printf("-n --nmin=N0\n");
  printf("-N --nmax=N1       Sieve for primes k*2^n+/-1 with N0 <= n < N1\n");
#ifndef PPS_TWIN
  printf("-R --riesel        Sieve for primes k*2^n-1 instead of +1.\n");
  printf("   --both          Sieve k*2^n+1 and k*2^n-1 in one pass.\n");
#endif
  printf("-d --device=N      Use GPU N instead of 0-threads\n");
}

/* Parameter rules and the factors file (pps/app.c:649-790): the rules are
 * ppsd_derive_params'; failures exit with the reference's code and text. */
void app_init(void) {
  int rc;
  int file_format = 0;
  print_factors = quiet_opt ? 0 : 1;
  if (input_filename != NULL) {
    /* pps/app.c:661-663, 188-347: the file's k, n ranges (and sign) fill
     * what the command line left open; its format is always taken from its
     * header */
    pps_input_scan scan;
    char err[300];
    if (pps_input_scan_file(input_filename, PPS_INPUT_TWIN, &scan, err, sizeof err)) {
      fprintf(stderr, "%s%s\n", bmprefix(), err);
      bexit(ERR_SCANF);
    }
    file_format = scan.format;
    if (kmin == 0 || kmax == 0 || nmin == 0 || nmax == 0) {
      if (scan.format == PPS_INPUT_ABCD) printf("Scanning ABCD file...\n");
#ifndef PPS_TWIN
      search_proth = scan.sign;
#endif
      printf("Found K's from %" PRIu64 " to %" PRIu64 ".\n", scan.k0, scan.k1);
      printf("Found N's from %u to %u.\n", scan.n0, scan.n1);
      if (kmin < scan.k0) kmin = scan.k0;
      if (kmax == 0 || kmax > scan.k1) kmax = scan.k1;
      if (nmin < scan.n0) nmin = scan.n0;
      if (nmax == 0 || nmax > scan.n1) nmax = scan.n1;
    }
  }
  memset(&params, 0, sizeof params);
  params.pmin = pmin;
  params.pmax = pmax;
  params.kmin = kmin;
  params.kmax = kmax;
  params.nmin = nmin;
  params.nmax = nmax;
  params.kstep = kstep;
  params.koffset = koffset;
  params.sign = search_proth;
  params.qmax = qmax_opt;
  rc = ppsd_derive_params(&params);
  if (rc) {
    char buf[200];
    snprintf(buf, sizeof buf, "%s\n", params.msg);
    bmsg(buf);
    bexit(rc == PPSD_ERR_FAILURE ? EXIT_FAILURE : rc);
  }
  kmin = params.kmin;
  kmax = params.kmax;
  nmin = params.nmin;
  nmax = params.nmax;
  printf("nstart=%u, nstep=%u\n", nmin, params.nstep_ref);
#ifdef PPS_TWIN
  {
    /* the twin sieve's windows overlap by one (pps/appcu.cu:177-180); the
     * factors do not depend on the window, so the walk keeps its own */
    const unsigned int tw = params.nstep_ref - 1;
    if (tw >= 32 && tw < 48 && tw != 32) printf("nstep changed to 32\n");
  }
#else
  if (params.nstep != params.nstep_ref) printf("nstep changed to %u\n", params.nstep);
#endif
  if (input_filename != NULL) {
    /* pps/app.c:726-738: one row per n in [nmin, nmax], one bit per k */
    char err[300];
    unsigned int count = 0;
    const uint64_t b0 = kmin / kstep, b1 = kmax / kstep;
    cand_row_bytes = (b1 - b0 + 8) / 8;
    cand_rows = nmax - nmin + 1;
    cand_bitmap = calloc((size_t)cand_rows, cand_row_bytes);
    if (!cand_bitmap) {
      bmsg("Cannot allocate the candidate bitmap\n");
      bexit(ERR_INSUFFICIENT_RESOURCE);
    }
    if (file_format == PPS_INPUT_ABCD) printf("Reading ABCD file.\n");
    if (pps_input_read_file(input_filename, file_format, PPS_INPUT_TWIN, kmin, kmax, nmin, nmax, kstep, koffset,
                            cand_bitmap, cand_row_bytes, &count, err, sizeof err)) {
      fprintf(stderr, "%s%s\n", bmprefix(), err);
      bexit(ERR_SCANF);
    }
    printf("Read %u terms from %s format input file `%s'\n", count,
           file_format == PPS_INPUT_ABCD ? "ABCD" : "NewPGen", input_filename);
  }
  if (factors_filename == NULL) factors_filename = FACTORS_FILENAME_DEFAULT;
  if ((factors_file = fopen(factors_filename, "a")) == NULL) {
    fprintf(stderr, "%sCannot open factors file `%s'\n", bmprefix(), factors_filename);
    bexit(ERR_FOPEN);
  }
  printf(APP_NAME " initialized: %" PRIu64 " <= k <= %" PRIu64 ", %u <= n < %u\n", kmin, kmax, nmin, nmax);
  fflush(stdout);
}

/* Per worker: open the device (pps/app.c:797-827 -> cuda_app_init).
 * Returns the batch size, 0 when the device cannot be used. */
unsigned int app_thread_init(int th) {
  const int dev = device_opt >= 0 ? device_opt : th;
  uint32_t batch = 0;
  int rc;
  if (th < 0 || th >= MAX_WORKERS) return 0;
  if (user_cthread_count) batch = user_cthread_count * 256u;
  rc = ppsd_open(dev, &params, batch, &ctxs[th][0]);
  if (rc) {
    fprintf(stderr, "%sGPU %d not usable: %s\n", bmprefix(), dev, ppsd_last_error());
    if (rc == PPSD_ERR_INVALID_PARAM) bexit(ERR_INVALID_PARAM);
    return 0;
  }
  if (ppsd_open(dev, &params, ppsd_batch_size(ctxs[th][0]), &ctxs[th][1])) ctxs[th][1] = NULL;
  for (int j = 0; j < 2 && cand_bitmap; j++) {
    if (!ctxs[th][j]) continue;
    rc = ppsd_set_candidates(ctxs[th][j], cand_bitmap, nmin, cand_rows, kmin / kstep, cand_row_bytes);
    if (rc) {
      fprintf(stderr, "%sGPU %d: candidate list: %s\n", bmprefix(), dev, ppsd_last_error());
      app_thread_fini(th);
      return 0;
    }
  }
  next_ctx[th] = 0;
  fcap[th] = 1024;
  fbuf[th] = malloc(fcap[th] * sizeof(ppsd_factor));
  fprintf(stderr, "%sThread %d: GPU %d, %u primes per batch\n", bmprefix(), th, dev, ppsd_batch_size(ctxs[th][0]));
  return ppsd_batch_size(ctxs[th][0]);
}

/* The context for thread th's next submission, once its previous batch
 * (two submissions back) is collected. */
static int take_ctx(int th, uint64_t *count, uint64_t *sum, int *rc) {
  const int j = ctxs[th][1] ? next_ctx[th] : 0;
  *rc = drain(th, j, count, sum);
  next_ctx[th] = j ^ 1;
  return j;
}

/* Collect both of thread th's contexts, the older submission first. */
static int drain_all(int th, uint64_t *count, uint64_t *sum) {
  for (int k = 0; k < 2; k++) {
    const int j = (next_ctx[th] + k) & 1;
    if (!ctxs[th][j]) continue;
    const int rc = drain(th, j, count, sum);
    if (rc) return rc;
  }
  return 0;
}

/* A batch of host primes (pps/app.c:1224-1250): collect the batch that was
 * submitted two calls ago, start this one, return without waiting; the
 * previous call's batch keeps walking meanwhile.  The device copies P, so
 * lastP is not needed to remember the batches in flight. */
void app_thread_fun(int th, const uint64_t *P, uint64_t *lastP, const unsigned int cthread_count) {
  int rc;
  const int j = take_ctx(th, NULL, NULL, &rc);
  if (rc) device_fail(th, "collect", rc);
  rc = ppsd_submit_primes(ctxs[th][j], P, cthread_count);
  if (rc) device_fail(th, "submit_primes", rc);
  pending[th][j] = 1;
}

/* Flush (pps/app.c:1257-1286): the last len primes, then drain everything,
 * so that a checkpoint may follow.  No padding is needed: a batch may be
 * shorter than cthread_count. */
void app_thread_fun1(int th, uint64_t *P, uint64_t *lastP, const unsigned int cthread_count, unsigned int len) {
  int rc;
  if (len > 0) app_thread_fun(th, P, lastP, len);
  rc = drain_all(th, NULL, NULL);
  if (rc) device_fail(th, "collect", rc);
}

/* All primes of [p0, p1) sieved on the device, in chunks that alternate
 * between the two contexts; returns with nothing in flight. */
int app_thread_range(int th, uint64_t p0, uint64_t p1, uint64_t *count, uint64_t *sum) {
  uint64_t p = p0;
  int rc;
  *count = 0;
  *sum = 0;
  while (p < p1) {
    uint64_t next = p1;
    const int j = take_ctx(th, count, sum, &rc);
    if (rc) return rc;
    rc = ppsd_submit_range(ctxs[th][j], p, p1, &next);
    if (rc) return rc;
    pending[th][j] = 1;
    p = next;
  }
  return drain_all(th, count, sum);
}

void app_thread_fini(int th) {
  for (int j = 0; j < 2; j++) {
    if (ctxs[th][j]) ppsd_close(ctxs[th][j]);
    ctxs[th][j] = NULL;
    pending[th][j] = 0;
  }
  free(fbuf[th]);
  fbuf[th] = NULL;
}

/* The application's checkpoint record, "nmin=..,nmax=..,factor_count=.."
 * (the line pps/app.c:1348-1352 writes).  It is accepted when it names this
 * search's exponent range and, if it counts factors already found, the
 * factors file they were appended to is not empty (pps/app.c:1301-1336
 * applies the same two conditions); otherwise the count restarts at 0. */
int app_read_checkpoint(FILE *fin) {
  static const char *const fields[] = {"nmin", "nmax", "factor_count"};
  uint64_t v[3];
  struct stat st;
  factor_count = 0;
  if (!ckpt_read_record(fin, fields, 3, v) || v[0] != nmin || v[1] != nmax || v[2] > UINT32_MAX) return 0;
  if (v[2] > 0 && (fstat(fileno(factors_file), &st) != 0 || st.st_size == 0)) {
    bmsg("Factors file missing.  Restarting test.\n");
    return 0;
  }
  factor_count = (unsigned int)v[2];
  return 1;
}

void app_write_checkpoint(FILE *fout) {
  fflush(factors_file);
  fprintf(fout, "nmin=%u,nmax=%u,factor_count=%u\n", nmin, nmax, factor_count);
}

void app_fini(uint64_t pstop) {
  (void)pstop;
  fclose(factors_file);
  factors_file = NULL;
  printf("Found %u factor%s\n", factor_count, (factor_count == 1) ? "" : "s");
}
