/* pps_app.h -- the PPSieve application hooks and the framework services they
 * use.  The hook list, signatures and calling contract are those of the
 * reference's pps/app.h:73-84 (called by main.c); the framework services are
 * those of main.h:49-60.  One hook is added, app_thread_range, for primes
 * generated on the device.
 */
#ifndef PPS_APP_H
#define PPS_APP_H

#include <stdint.h>
#include <stdio.h>

/* Built twice, as the reference builds ppsieve and tpsieve from one source
 * (pps/Makefile:155-162): with PPS_TWIN the program searches k*2^n+1 and
 * k*2^n-1 together for k = 6a+3 (the twin-prime sieve, pps/app.c:53-59). */
#ifdef PPS_TWIN
#define APP_PREFIX "tp"
#define KSTEP_DEFAULT 6
#define KOFFSET_DEFAULT 3
#else
#define APP_PREFIX "pp"
#define KSTEP_DEFAULT 2
#define KOFFSET_DEFAULT 1
#endif
#define APP_NAME APP_PREFIX "sieve"
#define APP_VERSION "mi355x-0.1"

#define CHECKPOINT_FILENAME APP_PREFIX "check%s.txt" /* pps/app.h:31 */
#define OLD_CHECKPOINT_FILENAME APP_PREFIX "checkpoint.txt"
#define CONFIG_FILENAME APP_PREFIX "config.txt"        /* pps/app.h:34 */
#define FACTORS_FILENAME_DEFAULT APP_PREFIX "factors.txt" /* pps/app.h:36 */

#define APP_SHORT_OPTS "k:K:n:N:i:f:b:d:m:M:s:R"
#define APP_LONG_OPTS                                                                                    \
  {"kmin", required_argument, 0, 'k'}, {"kmax", required_argument, 0, 'K'},                             \
      {"modulus", required_argument, 0, 'M'}, {"modshift", required_argument, 0, 's'},                  \
      {"nmin", required_argument, 0, 'n'}, {"nmax", required_argument, 0, 'N'},                         \
      {"input", required_argument, 0, 'i'}, {"factors", required_argument, 0, 'f'},                     \
      {"device", required_argument, 0, 'd'}, {"bitsatatime", required_argument, 0, 'b'},               \
      {"mthreads", required_argument, 0, 'm'}, {"riesel", no_argument, 0, 'R'},                         \
      {"both", no_argument, 0, 257},

/* Error codes, matching BOINC's (pps/app.h:64-70). */
#define ERR_FOPEN -108
#define ERR_SCANF -140
#define ERR_INVALID_PARAM -178
#define ERR_NEG -117
#define ERR_INSUFFICIENT_RESOURCE -198
#define ERR_NOT_IMPLEMENTED -150

/* hooks implemented by pps_app.c */
void app_banner(void);
int app_parse_option(int opt, char *arg, const char *source);
void app_help(void);
void app_init(void);
unsigned int app_thread_init(int th);
void app_thread_fun(int th, const uint64_t *P, uint64_t *lastP, const unsigned int cthread_count);
void app_thread_fun1(int th, uint64_t *P, uint64_t *lastP, const unsigned int cthread_count, unsigned int len);
/* New: every prime of [p0, p1) is generated and walked on the device; the
 * count and the sum (mod 2^64) of those primes are returned for the
 * framework's checkpoint and final count=/sum= line.  Returns 0, or a
 * negative code when the device failed. */
int app_thread_range(int th, uint64_t p0, uint64_t p1, uint64_t *count, uint64_t *sum);
void app_thread_fini(int th);
int app_read_checkpoint(FILE *f);
void app_write_checkpoint(FILE *f);
void app_fini(uint64_t pstop);

/* exported by the application (pps/app.h:93-96) */
extern unsigned int nmin, nmax;
extern int search_proth;
extern uint64_t kmax, kmin;

/* framework services (main.h:49-60) */
FILE *bfopen(const char *filename, const char *mode);
void bmsg(const char *msg);
char *bmprefix(void);
void bexit(int status);
extern int num_threads;
extern uint64_t pmin, pmax, pstart;
extern unsigned int quiet_opt;
extern unsigned int qmax_opt; /* sieve with odd primes <= qmax (main.c -Q) */

/* option values (putil.h) */
int parse_uint(unsigned int *result, const char *str, unsigned int lo, unsigned int hi);
int parse_uint64(uint64_t *result, const char *str, uint64_t lo, uint64_t hi);

#endif
