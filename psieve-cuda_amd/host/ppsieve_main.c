/* ppsieve_main.c -- the framework around the PPSieve application hooks:
 * options and ppconfig.txt, worker threads (one per GPU), checkpoints, status
 * reports and the closing statistics.
 *
 * Behaviour follows the reference's main.c (option names and meanings
 * main.c:375-650, checkpoint file main.c:305-372, final lines
 * main.c:1196-1257).  The prime supply is the difference: by default every
 * worker asks the device for all primes of a p-chunk (app_thread_range), so
 * no prime crosses PCIe; --host-primes instead runs a host segmented sieve
 * and feeds batches through app_thread_fun/app_thread_fun1 exactly as
 * main.c:736-832 does.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <getopt.h>
#include <inttypes.h>
#include <math.h>
#include <pthread.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include "pps_app.h"
#include "pps_ckpt.h"
#include "pps_hostsieve.h"

#define MAX_THREADS 32
#define PMIN_MIN 3
#define PMAX_MAX (UINT64_C(1) << 62)
#define QMAX_MAX (1U << 31)

int num_threads = 1;
uint64_t pmin = 0, pmax = 0, pstart = 0;
unsigned int quiet_opt = 0;
unsigned int qmax_opt = 0;

static char *pmin_str = NULL;
static char *checkpoint_filename = NULL;
static unsigned int report_opt = 60;       /* main.h:44 */
static unsigned int checkpoint_opt = 300;  /* main.h:45 */
static unsigned int priority_opt = 0;
static unsigned int blocksize_opt = 524288, chunksize_opt = 16384, blocks_opt = 2; /* accepted, unused */
static uint64_t chunk_span = 0; /* --chunk; 0: 2^28, or 2^30 with --host-primes */
static int host_primes = 0;
static int help_opt = 0;

static volatile sig_atomic_t stopping = 0;
static int got_sigint = 0, got_sigterm = 0;

/* ------------------------------------------------------------------ */
/* framework services (main.h)                                          */
/* ------------------------------------------------------------------ */
FILE *bfopen(const char *filename, const char *mode) { return fopen(filename, mode); }
void bmsg(const char *msg) { fprintf(stderr, "%s", msg); }
char *bmprefix(void) { return ""; }
void bexit(int status) { exit(status); }

static uint64_t now_usec(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000u + (uint64_t)ts.tv_nsec / 1000u;
}

static uint64_t cpu_usec(void) {
  struct rusage ru;
  getrusage(RUSAGE_SELF, &ru);
  return (uint64_t)(ru.ru_utime.tv_sec + ru.ru_stime.tv_sec) * 1000000u +
         (uint64_t)(ru.ru_utime.tv_usec + ru.ru_stime.tv_usec);
}

/* Integer option values with the suffix grammar of putil.c:86-168:
 * K M G T P (x10^3..), k m g t p (x2^10..), eN (x10^N), bN (x2^N).
 * Returns 0, -1 (cannot parse) or -2 (out of range). */
int parse_uint64(uint64_t *result, const char *str, uint64_t lo, uint64_t hi) {
  char *tail;
  uint64_t num;
  unsigned int e = 0;
  errno = 0;
  num = strtoull(str, &tail, 0);
  if (errno != 0 || num > hi) return -2;
  switch (*tail) {
    case 'P': e += 3; /* fall through */
    case 'T': e += 3; /* fall through */
    case 'G': e += 3; /* fall through */
    case 'M': e += 3; /* fall through */
    case 'K':
      e += 3;
      if (tail[1] != '\0') return -1;
      for (; e > 0; e -= 3) {
        if (num > hi / 1000) return -2;
        num *= 1000;
      }
      break;
    case 'e':
    case 'E':
      e = (unsigned int)strtoul(tail + 1, &tail, 0);
      if (errno != 0) return -2;
      if (*tail != '\0') return -1;
      while (e-- > 0) {
        if (num > hi / 10) return -2;
        num *= 10;
      }
      break;
    case 'p': e += 10; /* fall through */
    case 't': e += 10; /* fall through */
    case 'g': e += 10; /* fall through */
    case 'm': e += 10; /* fall through */
    case 'k':
      e += 10;
      if (tail[1] != '\0') return -1;
      if (num > (hi >> e)) return -2;
      num <<= e;
      break;
    case 'b':
    case 'B':
      e = (unsigned int)strtoul(tail + 1, &tail, 0);
      if (errno != 0) return -2;
      if (*tail != '\0') return -1;
      while (e-- > 0) {
        if (num > (hi >> 1)) return -2;
        num <<= 1;
      }
      break;
    case '\0': break;
    default: return -1;
  }
  if (num < lo) return -2;
  *result = num;
  return 0;
}

int parse_uint(unsigned int *result, const char *str, unsigned int lo, unsigned int hi) {
  uint64_t v;
  int st = parse_uint64(&v, str, lo, hi);
  if (st == 0) *result = (unsigned int)v;
  return st;
}

/* ------------------------------------------------------------------ */
/* options                                                              */
/* ------------------------------------------------------------------ */
static const char *short_opts = "p:P:Q:B:C:c:r:t:z:qh" APP_SHORT_OPTS;
static const struct option long_opts[] = {
    {"pmin", required_argument, 0, 'p'},
    {"pmax", required_argument, 0, 'P'},
    {"qmax", required_argument, 0, 'Q'},
    {"blocksize", required_argument, 0, 'B'},
    {"chunksize", required_argument, 0, 'C'},
    {"blocks", required_argument, 0, 256},
    {"checkpoint", required_argument, 0, 'c'},
    {"report", required_argument, 0, 'r'},
    {"nthreads", required_argument, 0, 't'},
    {"priority", required_argument, 0, 'z'},
    {"help", no_argument, 0, 'h'},
    {"quiet", no_argument, 0, 'q'},
    {"host-primes", no_argument, 0, 258},
    {"chunk", required_argument, 0, 259},
    APP_LONG_OPTS{0, 0, 0, 0}};

static void help(void) {
  printf("-p --pmin=P0       Sieve start: 3 <= P0 <= p (default P0=3)\n");
  printf("-P --pmax=P1       Sieve end: p < P1 <= 2^62 (default P1=P0+10^9)\n");
  printf("-Q --qmax=Q1       Sieve only with odd primes q < Q1 <= 2^31\n");
  printf("-c --checkpoint=N  Checkpoint every N seconds (default N=%u)\n", checkpoint_opt);
  printf("-q --quiet         Don't print factors to screen\n");
  printf("-r --report=N      Report status every N seconds (default N=%u)\n", report_opt);
  printf("-t --nthreads=N    Start N threads on N GPUs (default N=1)\n");
  printf("-z --priority=N    Set process priority to nice N or {idle,low,normal}\n");
  printf("   --host-primes   Sieve primes on the host and pass them in batches\n");
  printf("   --chunk=N       p-range handed to a worker at a time (default 2^28, 2^30 with --host-primes)\n");
  printf("-h --help          Print this help\n");
}

static int parse_option(int opt, char *arg, const char *source) {
  int status = 0;
  switch (opt) {
    case 'p':
      free(pmin_str);
      pmin_str = strdup(arg);
      status = parse_uint64(&pmin, arg, PMIN_MIN, PMAX_MAX - 1);
      break;
    case 'P': status = parse_uint64(&pmax, arg, 4, PMAX_MAX); break;
    case 'Q': status = parse_uint(&qmax_opt, arg, 3, QMAX_MAX); break;
    case 'B': status = parse_uint(&blocksize_opt, arg, 1024, 1 << 27); break;
    case 'C': status = parse_uint(&chunksize_opt, arg, 16, 1 << 19); break;
    case 256: status = parse_uint(&blocks_opt, arg, 2, 4); break;
    case 'c': status = parse_uint(&checkpoint_opt, arg, 0, UINT32_MAX); break;
    case 'r': status = parse_uint(&report_opt, arg, 0, UINT32_MAX); break;
    case 't': status = parse_uint((unsigned int *)&num_threads, arg, 1, MAX_THREADS); break;
    case 'z':
      if (strcmp(arg, "idle") == 0)
        priority_opt = 19;
      else if (strcmp(arg, "low") == 0)
        priority_opt = 10;
      else if (strcmp(arg, "normal") == 0)
        priority_opt = 0;
      else if (strcmp(arg, "none") == 0)
        priority_opt = (unsigned int)-1;
      else
        status = parse_uint(&priority_opt, arg, 0, 19);
      priority_opt++;
      break;
    case 'h': help_opt = 1; break;
    case '?': status = -3; break;
    case 'q': quiet_opt = 1; break;
    case 258: host_primes = 1; break;
    case 259: status = parse_uint64(&chunk_span, arg, 1 << 20, PMAX_MAX); break;
    default: status = app_parse_option(opt, arg, source); break;
  }
  return status;
}

static void process_args(int argc, char *argv[]) {
  int ind = -1, opt;
  while ((opt = getopt_long(argc, argv, short_opts, long_opts, &ind)) != -1) {
    switch (parse_option(opt, optarg, NULL)) {
      case 0: ind = -1; break;
      case -1:
        if (ind == -1)
          fprintf(stderr, "%s%s: invalid argument -%c %s\n", bmprefix(), argv[0], opt, optarg);
        else
          fprintf(stderr, "%s%s: invalid argument --%s %s\n", bmprefix(), argv[0], long_opts[ind].name, optarg);
        exit(EXIT_FAILURE);
      case -2:
        if (ind == -1)
          fprintf(stderr, "%s%s: out of range argument -%c %s\n", bmprefix(), argv[0], opt, optarg);
        else
          fprintf(stderr, "%s%s: out of range argument --%s %s\n", bmprefix(), argv[0], long_opts[ind].name,
                  optarg);
        exit(EXIT_FAILURE);
      default: exit(EXIT_FAILURE);
    }
  }
  while (optind < argc) {
    if (parse_option(0, argv[optind], NULL) != 0) {
      fprintf(stderr, "%s%s: invalid non-option argument %s\n", bmprefix(), argv[0], argv[optind]);
      exit(EXIT_FAILURE);
    }
    optind++;
  }
  if (help_opt) {
    help();
    app_help();
    exit(EXIT_SUCCESS);
  }
}

/* One option of ppconfig.txt by its long name; arg is NULL when the line
 * has no value.  The messages are the reference's (main.c:597-646). */
static void apply_config_option(const char *fn, const char *name, char *arg) {
  const struct option *o = long_opts;
  while (o->name && strcmp(o->name, name) != 0) o++;
  if (!o->name) {
    fprintf(stderr, "%s%s: unrecognised option `%s'\n", bmprefix(), fn, name);
    exit(EXIT_FAILURE);
  }
  if (o->has_arg == no_argument && arg) {
    fprintf(stderr, "%s%s: option `%s' doesn't allow an argument\n", bmprefix(), fn, name);
    exit(EXIT_FAILURE);
  }
  if (o->has_arg == required_argument && !arg) {
    fprintf(stderr, "%s%s: option `%s' requires an argument\n", bmprefix(), fn, name);
    exit(EXIT_FAILURE);
  }
  const int st = parse_option(o->val, arg, fn);
  if (st == 0) return;
  fprintf(stderr, "%s%s: %s argument %s %s\n", bmprefix(), fn,
          st == -1 ? "invalid" : st == -2 ? "out of range" : "weird", name, arg ? arg : "");
  exit(EXIT_FAILURE);
}

/* ppconfig.txt (read before the command line, which overrides it): one
 * option per line by its long name -- "name=value", "name value" or a bare
 * "name" -- with blank lines and '#' comment lines skipped. */
static void read_config_file(const char *fn) {
  static const char seps[] = "= \t\r\n\v";
  char line[256];
  FILE *f = bfopen(fn, "r");
  if (!f) return;
  while (fgets(line, sizeof line, f)) {
    char *name = line + strspn(line, " \t\r\n\v");
    if (*name == '\0' || *name == '#') continue;
    char *name_end = name + strcspn(name, seps);
    char *val = name_end + strspn(name_end, seps);
    val[strcspn(val, " \t\r\n\v")] = '\0';
    *name_end = '\0';
    apply_config_option(fn, name, *val ? val : NULL);
  }
  fclose(f);
}

/* ------------------------------------------------------------------ */
/* work distribution and checkpoints                                    */
/* ------------------------------------------------------------------ */
typedef struct {
  uint64_t p0, p1, count, sum;
  int done;
} chunk_t;

static pthread_mutex_t mtx = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t cond = PTHREAD_COND_INITIALIZER;
static chunk_t *chunks;
static size_t nchunks, next_chunk_ix, frontier_ix;
static uint64_t cand_count, cand_sum; /* primes below the frontier (incl. resumed) */
static int checkpoint_request, idle_workers, live_workers, unusable_workers, failed;

/* Advance the frontier over completed chunks; caller holds mtx. */
static void advance_frontier(void) {
  while (frontier_ix < nchunks && chunks[frontier_ix].done) {
    cand_count += chunks[frontier_ix].count;
    cand_sum += chunks[frontier_ix].sum;
    frontier_ix++;
  }
}

static uint64_t progress_p(void) { return frontier_ix < nchunks ? chunks[frontier_ix].p0 : pmax; }

static void write_checkpoint(uint64_t p) {
  FILE *fout = fopen(checkpoint_filename, "w");
  if (!fout) return;
  fprintf(fout, "pmin=%" PRIu64 ",p=%" PRIu64 ",count=%" PRIu64 ",sum=0x%016" PRIx64 ",checksum=0x%016" PRIx64 "\n",
          pmin, p, cand_count, cand_sum, pmin + p + cand_count + cand_sum);
  app_write_checkpoint(fout);
  fclose(fout);
}

/* The checkpoint (format of main.c:305-330) is taken when its first record
 * names this pmin and a resume point inside (pmin, pmax), the application
 * accepts the second record and the checksum -- the wrapping sum of pmin,
 * p, count and sum -- holds (the acceptance rule of main.c:332-372).  A
 * file left under the reference's older name is read when the current one
 * is absent. */
static uint64_t read_checkpoint(void) {
  static const char *const fields[] = {"pmin", "p", "count", "sum", "checksum"};
  enum { F_PMIN, F_P, F_COUNT, F_SUM, F_CHECK, NF };
  uint64_t v[NF];
  const char *cpf = checkpoint_filename;
  FILE *fin = fopen(cpf, "r");
  int ok;
  if (!fin) {
    cpf = OLD_CHECKPOINT_FILENAME;
    fin = fopen(cpf, "r");
    if (!fin) return pmin;
  }
  ok = ckpt_read_record(fin, fields, NF, v) && v[F_PMIN] == pmin && v[F_P] > pmin && v[F_P] < pmax;
  ok = ok && app_read_checkpoint(fin);
  ok = ok && v[F_PMIN] + v[F_P] + v[F_COUNT] + v[F_SUM] == v[F_CHECK];
  fclose(fin);
  if (!ok) {
    fprintf(stderr, "%sIgnoring invalid checkpoint in %s\n", bmprefix(), cpf);
    return pmin;
  }
  fprintf(stderr, "%sResuming from checkpoint p=%" PRIu64 " in %s\n", bmprefix(), v[F_P], cpf);
  cand_count = v[F_COUNT];
  cand_sum = v[F_SUM];
  return v[F_P];
}

/* Next chunk for a worker, or NULL when done/stopping; honours checkpoint
 * requests by parking the worker until the checkpoint is written. */
static chunk_t *take_chunk(void) {
  chunk_t *c = NULL;
  pthread_mutex_lock(&mtx);
  while (checkpoint_request && !stopping) {
    idle_workers++;
    pthread_cond_broadcast(&cond);
    while (checkpoint_request && !stopping) pthread_cond_wait(&cond, &mtx);
    idle_workers--;
  }
  if (!stopping && !failed && next_chunk_ix < nchunks) c = &chunks[next_chunk_ix++];
  pthread_mutex_unlock(&mtx);
  return c;
}

static void finish_chunk(chunk_t *c, uint64_t count, uint64_t sum) {
  pthread_mutex_lock(&mtx);
  c->count = count;
  c->sum = sum;
  c->done = 1;
  advance_frontier();
  pthread_cond_broadcast(&cond);
  pthread_mutex_unlock(&mtx);
}

/* odd primes <= lim (host sieve for --host-primes) */
static uint32_t *host_sieve_primes(uint32_t lim, uint32_t *n) {
  uint8_t *comp = calloc((size_t)lim / 2 + 1, 1);
  uint32_t *out, cnt = 0;
  for (uint64_t i = 3; i * i <= lim; i += 2)
    if (!comp[i / 2])
      for (uint64_t m = i * i; m <= lim; m += 2 * i) comp[m / 2] = 1;
  for (uint64_t i = 3; i <= lim; i += 2) cnt += !comp[i / 2];
  out = malloc(sizeof(uint32_t) * (cnt + 1));
  cnt = 0;
  for (uint64_t i = 3; i <= lim; i += 2)
    if (!comp[i / 2]) out[cnt++] = (uint32_t)i;
  free(comp);
  *n = cnt;
  return out;
}

static uint32_t *hq;
static uint32_t nhq;

/* --host-primes: primes of a chunk gathered into batches of cthread_count
 * (main.c:736-832) and handed to the application as they fill */
struct batcher {
  int th;
  uint64_t *P, *lastP;
  unsigned int plen, cap;
  uint64_t count, sum;
};

static void batch_primes(void *arg, const uint64_t *p, size_t n) {
  struct batcher *b = arg;
  for (size_t i = 0; i < n; i++) {
    b->count++;
    b->sum += p[i];
    b->P[b->plen++] = p[i];
    if (b->plen == b->cap) {
      app_thread_fun(b->th, b->P, b->lastP, b->cap);
      b->plen = 0;
    }
  }
}

static void *thread_fun(void *arg) {
  const int th = (int)(long)arg;
  unsigned int cthread_count = app_thread_init(th);
  uint64_t *P = NULL, *lastP = NULL;
  pps_hostsieve *hs = NULL;
  fprintf(stderr, "%sThread %d starting\n", bmprefix(), th);
  if (cthread_count == 0) {
    pthread_mutex_lock(&mtx);
    unusable_workers++;
    live_workers--;
    pthread_cond_broadcast(&cond);
    pthread_mutex_unlock(&mtx);
    return NULL;
  }
  if (host_primes) {
    P = malloc(sizeof(uint64_t) * cthread_count);
    lastP = malloc(sizeof(uint64_t) * cthread_count);
    /* the host's cores shared among the GPU workers, at most 16 sieve threads each */
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    long per = (ncpu > 0 ? ncpu : 1) / (num_threads > 0 ? num_threads : 1);
    hs = pps_hostsieve_new(hq, nhq, (int)(per < 1 ? 1 : per > 16 ? 16 : per));
    if (!P || !lastP || !hs) {
      fprintf(stderr, "%sThread %d: out of memory for the host sieve\n", bmprefix(), th);
      exit(EXIT_FAILURE);
    }
  }
  for (chunk_t *c; (c = take_chunk()) != NULL;) {
    uint64_t count = 0, sum = 0;
    if (!host_primes) {
      int rc = app_thread_range(th, c->p0, c->p1, &count, &sum);
      if (rc) {
        fprintf(stderr, "%sThread %d: device error %d\n", bmprefix(), th, rc);
        pthread_mutex_lock(&mtx);
        failed = 1;
        pthread_cond_broadcast(&cond);
        pthread_mutex_unlock(&mtx);
        break;
      }
    } else {
      struct batcher b = {th, P, lastP, 0, cthread_count, 0, 0};
      if (pps_hostsieve_range(hs, c->p0, c->p1, batch_primes, &b)) {
        fprintf(stderr, "%sThread %d: out of memory in the host sieve\n", bmprefix(), th);
        exit(EXIT_FAILURE);
      }
      app_thread_fun1(th, P, lastP, cthread_count, b.plen);
      count = b.count;
      sum = b.sum;
    }
    finish_chunk(c, count, sum);
  }
  app_thread_fini(th);
  free(P);
  free(lastP);
  pps_hostsieve_free(hs);
  pthread_mutex_lock(&mtx);
  live_workers--;
  pthread_cond_broadcast(&cond);
  pthread_mutex_unlock(&mtx);
  return NULL;
}

static void handle_signal(int signum) {
  if (signum == SIGINT) got_sigint = 1;
  if (signum == SIGTERM) got_sigterm = 1;
  stopping = 1;
}

/* The status line (main.c:246-300): the frontier p, the p-range rate since
 * the last report (K/M units, about four significant digits), the CPU cores
 * the process kept busy meanwhile, the fraction of [pmin, pmax) done, and
 * the finish time extrapolated from the progress since this run started at
 * p_resume (strftime "ETA %d %b %H:%M", empty before any progress). */
static void report_status(uint64_t progress, uint64_t p_resume, time_t run_started, uint64_t dt_usec, uint64_t dp,
                          uint64_t dcpu_usec) {
  static const char *const units[] = {"", "K", "M"};
  double rate = (double)dp / (double)(dt_usec ? dt_usec : 1) * 1e6; /* p per second */
  int u = 0, digits;
  char eta[32] = "";
  while (u < 2 && rate >= 1000.0) rate /= 1000.0, u++;
  digits = rate < 10.0 ? 3 : rate < 100.0 ? 2 : rate < 1000.0 ? 1 : 0;
  if (progress > p_resume && pmax > p_resume) {
    const double frac = (double)(progress - p_resume) / (double)(pmax - p_resume);
    const time_t finish = run_started + (time_t)((double)(time(NULL) - run_started) / frac);
    struct tm tmv;
    if (!localtime_r(&finish, &tmv) || !strftime(eta, sizeof eta, "ETA %d %b %H:%M", &tmv)) eta[0] = '\0';
  }
  printf("p=%" PRIu64 ", %.*f%s p/sec, %.2f CPU cores, %.1f%% done. %s  ", progress, digits, rate, units[u],
         (double)dcpu_usec / (double)(dt_usec ? dt_usec : 1), 100.0 * (double)(progress - pmin) / (double)(pmax - pmin),
         eta);
  if (quiet_opt) {
    putchar('\r');
    fflush(stdout);
  } else
    putchar('\n');
}

int main(int argc, char *argv[]) {
  const uint64_t program_start = now_usec();
  uint64_t sieve_start, cpu_start, pstop, last_ckpt, last_rep, last_rep_p, last_rep_cpu;
  time_t run_started;
  pthread_t tid[MAX_THREADS];
  int th, process_ret = EXIT_SUCCESS;

  setvbuf(stdout, NULL, _IOLBF, 0);
  app_banner();
  read_config_file(CONFIG_FILENAME);
  process_args(argc, argv);

  if (pmin < PMIN_MIN) pmin = PMIN_MIN;
  if (pmax > PMAX_MAX) pmax = PMAX_MAX;
  if (pmin >= pmax) {
    if (pmax == 0 && pmin < PMAX_MAX - 1000000000) {
      fprintf(stderr, "%spmax not specified, using default pmax = pmin + 1e9\n", bmprefix());
      pmax = pmin + 1000000000;
    } else {
      fprintf(stderr, "%sOption out of range: pmax must be greater than pmin\n", bmprefix());
      exit(EXIT_FAILURE);
    }
  }
  if (asprintf(&checkpoint_filename, CHECKPOINT_FILENAME, pmin_str ? pmin_str : "") < 0) exit(EXIT_FAILURE);
  if (priority_opt) setpriority(PRIO_PROCESS, 0, (int)priority_opt - 1);
  if (qmax_opt && (uint64_t)qmax_opt * qmax_opt > pmax) qmax_opt = 0; /* default: all <= sqrt(pmax) */

  app_init();
  fprintf(stderr, "%sSieve started: %" PRIu64 " <= p < %" PRIu64 "\n", bmprefix(), pmin, pmax);

  pstart = read_checkpoint() | 1;

  if (host_primes) {
    uint64_t lim = (uint64_t)sqrt((double)(pmax - 1));
    while (lim * lim > pmax - 1) lim--;
    while ((lim + 1) * (lim + 1) <= pmax - 1) lim++;
    if (qmax_opt && qmax_opt < lim) lim = qmax_opt < 61 ? 61 : qmax_opt;
    hq = host_sieve_primes((uint32_t)lim, &nhq);
  }

  /* chunks [pstart, pmax) */
  /* host sieving spreads a chunk over several threads, which pays on longer chunks */
  if (!chunk_span) chunk_span = host_primes ? UINT64_C(1) << 30 : UINT64_C(1) << 28;
  nchunks = (size_t)((pmax - pstart + chunk_span - 1) / chunk_span);
  if (pstart >= pmax) nchunks = 0;
  chunks = calloc(nchunks + 1, sizeof(chunk_t));
  for (size_t i = 0; i < nchunks; i++) {
    chunks[i].p0 = pstart + i * chunk_span;
    chunks[i].p1 = (i + 1 == nchunks) ? pmax : chunks[i].p0 + chunk_span;
  }

  signal(SIGINT, handle_signal);
  signal(SIGTERM, handle_signal);
  signal(SIGHUP, handle_signal);

  sieve_start = now_usec();
  run_started = time(NULL);
  cpu_start = cpu_usec();
  last_ckpt = last_rep = sieve_start;
  last_rep_p = pstart;
  last_rep_cpu = cpu_start;
  live_workers = num_threads;
  for (th = 0; th < num_threads; th++) pthread_create(&tid[th], NULL, thread_fun, (void *)(long)th);

  pthread_mutex_lock(&mtx);
  while (live_workers > 0) {
    struct timespec ts;
    uint64_t t = now_usec(), wake;
    if (report_opt && t >= last_rep + (uint64_t)report_opt * 1000000u) {
      uint64_t cpu = cpu_usec();
      report_status(progress_p(), pstart, run_started, t - last_rep, progress_p() - last_rep_p, cpu - last_rep_cpu);
      last_rep = t;
      last_rep_p = progress_p();
      last_rep_cpu = cpu;
    }
    if (checkpoint_opt && !stopping && t >= last_ckpt + (uint64_t)checkpoint_opt * 1000000u) {
      /* park every worker between chunks (main.c:1063-1098) */
      checkpoint_request = 1;
      while (idle_workers < live_workers && !stopping) pthread_cond_wait(&cond, &mtx);
      write_checkpoint(progress_p());
      checkpoint_request = 0;
      pthread_cond_broadcast(&cond);
      last_ckpt = now_usec();
    }
    wake = t + 200000;
    clock_gettime(CLOCK_REALTIME, &ts);
    ts.tv_nsec += 200000000;
    if (ts.tv_nsec >= 1000000000) ts.tv_sec++, ts.tv_nsec -= 1000000000;
    (void)wake;
    pthread_cond_timedwait(&cond, &mtx, &ts);
    if (stopping) pthread_cond_broadcast(&cond);
  }
  pthread_mutex_unlock(&mtx);
  for (th = 0; th < num_threads; th++) pthread_join(tid[th], NULL);
  if (failed || unusable_workers == num_threads) process_ret = EXIT_FAILURE;

  /* a failed run leaves the last checkpoint as it was (main.c:1196-1204) */
  pstop = progress_p();
  if (pstop > pmax) pstop = pmax;
  if (process_ret == EXIT_SUCCESS)
    write_checkpoint(pstop);
  else
    pstop = pstart;
  if (pstop >= pmax) {
    fprintf(stderr, "\n%sSieve complete: %" PRIu64 " <= p < %" PRIu64 "\n", bmprefix(), pmin, pmax);
    remove(checkpoint_filename);
  } else
    fprintf(stderr, "\n%sSieve incomplete: %" PRIu64 " <= p < %" PRIu64 "\n", bmprefix(), pmin, pstop);
  app_fini(pstop);
  if (process_ret == EXIT_SUCCESS)
    fprintf(stderr, "%scount=%" PRIu64 ",sum=0x%016" PRIx64 "\n", bmprefix(), cand_count, cand_sum);
  {
    const uint64_t stop = now_usec(), cpu = cpu_usec();
    fprintf(stderr, "%sElapsed time: %.2f sec. (%.2f init + %.2f sieve) at %.0f p/sec.\n", bmprefix(),
            (stop - program_start) / 1e6, (sieve_start - program_start) / 1e6, (stop - sieve_start) / 1e6,
            (double)(pstop - pstart) / (double)(stop - sieve_start + 1) * 1e6);
    fprintf(stderr, "%sProcessor time: %.2f sec. at %.0f p/sec.\n", bmprefix(), cpu / 1e6,
            (double)(pstop - pstart) / (double)(cpu - cpu_start + 1) * 1e6);
  }
  if (got_sigint) {
    signal(SIGINT, SIG_DFL);
    raise(SIGINT);
  }
  if (got_sigterm) {
    signal(SIGTERM, SIG_DFL);
    raise(SIGTERM);
  }
  return process_ret;
}
