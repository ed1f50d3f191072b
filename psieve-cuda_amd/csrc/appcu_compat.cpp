// appcu_compat.cpp -- pps/appcu.h implemented on libppsd: the entry points an
// unmodified PPSieve host (main.c + pps/app.c) calls, so that libppsd.so
// replaces pps/appcu.cu at link time.
//
// The reference device layer only flags, per prime, which n-subsections
// hold a candidate (one byte per prime) and lets pps/app.c re-walk them on
// the CPU (pps/app.c:1197-1219).  Here the device returns the exact divisors
// (ppsd_collect_raw) and get_factors_found turns each into the flag bit of
// the subsection holding its n, so the host's re-walk finds exactly them.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "../../include/ppsd.h"
#include "../../include/ppsd_appcu.h"

extern "C" {
// application and framework symbols (pps/app.h:93-96, main.h:49-60), bound
// to the host program at load time
extern uint64_t pmin __attribute__((weak));
extern uint64_t pmax __attribute__((weak));
extern uint64_t kmin __attribute__((weak));
extern uint64_t kmax __attribute__((weak));
extern unsigned int nmin __attribute__((weak));
extern unsigned int nmax __attribute__((weak));
extern int search_proth __attribute__((weak));
extern struct n_subsection *thread_subsections __attribute__((weak));
void bmsg(const char *msg) __attribute__((weak));
void bexit(int status) __attribute__((weak));
char *bmprefix(void) __attribute__((weak));

unsigned int ld_nstep;
int ld_bbits;
uint64_t ld_r0;
}

namespace {

constexpr int kMaxWorkers = 32;  // main.h:24
constexpr int kErrInvalidParam = -178;

struct Worker {
  ppsd_ctx *ctx = nullptr;
  std::vector<uint64_t> P;                          // batch in flight
  std::vector<std::pair<uint64_t, uint32_t>> byp;  // (prime, index), sorted
  std::vector<ppsd_factor> hits;
};

Worker g_workers[kMaxWorkers];
thread_local int t_th = -1;

bool host_present() {
  return &pmin && &pmax && &kmin && &kmax && &nmin && &nmax && &search_proth && &thread_subsections && bexit &&
         bmsg && bmprefix;
}

[[noreturn]] void fail(int status) {
  bexit(status);
  exit(status);
}

int lg2(uint64_t v) {
  int r = 0;
  while (v >>= 1) ++r;
  return r;
}

// The reference builds its twin-prime sieve from the same sources with
// SEARCH_TWIN (pps/Makefile:155-162), which changes the device layer too.
// One library serves both: the twin host sets PPSD_APPCU_TWIN=1.
bool twin_mode() {
  const char *e = getenv("PPSD_APPCU_TWIN");
  return e && *e && *e != '0';
}

// Measurement mode, PPSD_APPCU_FLAG_ALL=1: get_factors_found flags every
// n-subsection of every prime, so that the host re-walks the whole n-range
// of each prime on the CPU with its own test_one_p (pps/app.c:1037-1121,
// 1197-1219).  That times the reference's CPU walk through its unmodified
// code; the results are unchanged (a re-walk finds exactly the divisors).
bool flag_all_mode() {
  const char *e = getenv("PPSD_APPCU_FLAG_ALL");
  return e && *e && *e != '0';
}

void device_error(const char *where) {
  fprintf(stderr, "Cuda error: %s: %s\n", where, ppsd_last_error());  // checkCUDAErr, appcu.cu:124-131
  cuda_finalize();
  fail(EXIT_FAILURE);
}

}  // namespace

extern "C" {

void cuda_init(void) {
  if (!host_present()) {
    fprintf(stderr, "cuda_init: the PPSieve host symbols are not linked\n");
    return;
  }
  if (ld_nstep > nmax - nmin + 1) ld_nstep = nmax - nmin + 1;
  if (((uint64_t)1 << (64 - ld_nstep)) > pmin) {
    uint64_t pm = pmin, pm1 = (uint64_t)1 << (64 - ld_nstep);
    unsigned int s = ld_nstep;
    bmsg("Error: pmin is not large enough (or nmax is close to nmin).\n");
    while (((uint64_t)1 << (64 - s)) > pm) {
      pm *= 2;
      ++s;
    }
    if (pm1 < pm) pm = pm1;
    fprintf(stderr, "This program will work by the time pmin == %lu.\n", (unsigned long)pm);
    fail(kErrInvalidParam);
  }
  // the twin sieve checks k*2^n+1 and k*2^n-1 with windows that overlap by
  // one (pps/appcu.cu:177-180)
  if (twin_mode()) ld_nstep--;
  if (ld_nstep >= 32 && ld_nstep < 48 && ((uint64_t)1 << 32) <= pmin) {
    if (ld_nstep != 32) printf("nstep changed to 32\n");
    ld_nstep = 32;
  } else {
    printf("Didn't change nstep from %u\n", ld_nstep);
  }
  ld_bbits = lg2(nmin);
  if (ld_bbits < 6) {
    fprintf(stderr, "%sError: nmin too small at %d (must be at least 64).\n", bmprefix(), nmin);
    fail(kErrInvalidParam);
  }
  ld_r0 = (uint64_t)1 << (64 - (nmin >> (ld_bbits - 5)));
  ld_bbits -= 6;
}

unsigned int cuda_app_init(int gpuno, int th, unsigned int cthread_count) {
  if (!host_present() || th < 0 || th >= kMaxWorkers) return 0;
  if (pmin <= ((uint64_t)1 << 32)) {
    bmsg("Error: PMin is too small, <= 2^32!\n");
    fail(kErrInvalidParam);
  }
  Worker &w = g_workers[th];
  t_th = th;
  uint32_t batch = 0;
  if (cthread_count) {  // -m: per compute unit, whole workgroups (appcu.cu:255-266)
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, gpuno) == hipSuccess) cus = prop.multiProcessorCount;
    cthread_count = cthread_count < 128 ? cthread_count * 128 : cthread_count - cthread_count % 128;
    batch = cthread_count * (uint32_t)cus;
  }
  ppsd_params p;
  memset(&p, 0, sizeof p);
  p.pmin = pmin;
  p.pmax = pmax;
  p.kmin = kmin;
  p.kmax = kmax;
  p.nmin = nmin;
  p.nmax = nmax;
  p.kstep = 2;  // raw divisors need only k odd; pps/app.c applies its own -M/-s rule
  p.koffset = 1;
  p.sign = twin_mode() ? 0 : search_proth;  // twin: both forms
  p.nstep_ref = ld_nstep;
  p.nstep = ld_nstep > 32 ? 32 : ld_nstep;
  if (ppsd_open(gpuno, &p, batch, &w.ctx) != 0) {
    fprintf(stderr, "%sGPU %d not usable: %s\n", bmprefix(), gpuno, ppsd_last_error());
    return 0;
  }
  fprintf(stderr, "%sDetected GPU %d (libppsd)\n", bmprefix(), gpuno);
  // Eight n-subsections [start[j+1], start[j]), boundaries on the window grid
  // nmin + m*ld_nstep so that re-walks of adjacent subsections never overlap;
  // start[8] = nmin, start[0] = nmax (appcu.cu:366-393 lays them out alike).
  unsigned int *start = thread_subsections[th].n_subsection_start;
  start[8] = nmin;
  for (int j = 7; j >= 1; --j) {
    const uint64_t target = nmin + ((uint64_t)(nmax - nmin) * (8 - j)) / 8;
    uint64_t v = nmin + (target - nmin + ld_nstep - 1) / ld_nstep * ld_nstep;
    if (v > nmax) v = nmax;
    if (v < start[j + 1]) v = start[j + 1];
    start[j] = (unsigned int)v;
  }
  start[0] = nmax;
  thread_subsections[th].first_n_subsection = &start[7];
  return ppsd_batch_size(w.ctx);
}

void check_ns(const uint64_t *P, const unsigned int cthread_count, const int th) {
  if (th < 0 || th >= kMaxWorkers || !g_workers[th].ctx) return;
  Worker &w = g_workers[th];
  t_th = th;
  if (P) w.P.assign(P, P + cthread_count);
  if (ppsd_submit_primes(w.ctx, w.P.data(), cthread_count) != 0) device_error("check_ns");
  w.byp.resize(w.P.size());
  for (uint32_t i = 0; i < w.P.size(); ++i) w.byp[i] = {w.P[i], i};
  std::sort(w.byp.begin(), w.byp.end());
}

void get_factors_found(unsigned char *factor_found, const unsigned int cthread_count, const uint64_t start_t,
                       int *check_ns_delay) {
  (void)start_t;
  (void)check_ns_delay;
  const int th = t_th;
  if (th < 0 || !g_workers[th].ctx) return;
  Worker &w = g_workers[th];
  uint64_t n = 0;
  memset(factor_found, 0, cthread_count);
  if (ppsd_collect_raw(w.ctx, nullptr, 0, &n, nullptr, nullptr) != 0) device_error("get_factors_found");
  w.hits.resize(n);
  if (n && ppsd_collect_raw(w.ctx, w.hits.data(), n, &n, nullptr, nullptr) != 0) device_error("get_factors_found");
  if (flag_all_mode()) {
    // every prime of the batch; a flushed batch is padded with multiples of
    // 3 (pps/app.c:1261-1266), which are left out
    for (unsigned int i = 0; i < cthread_count && i < w.P.size(); ++i) factor_found[i] = w.P[i] % 3 ? 0xff : 0;
    return;
  }
  const unsigned int *start = thread_subsections[th].n_subsection_start;
  for (const ppsd_factor &h : w.hits) {
    int j = 0;
    while (j < 8 && !(start[j + 1] <= h.n && h.n < start[j])) ++j;
    if (j == 8) continue;
    auto range = std::equal_range(w.byp.begin(), w.byp.end(), std::make_pair(h.p, 0u),
                                  [](const std::pair<uint64_t, uint32_t> &a, const std::pair<uint64_t, uint32_t> &b) {
                                    return a.first < b.first;
                                  });
    for (auto it = range.first; it != range.second; ++it)
      if (it->second < cthread_count) factor_found[it->second] |= (unsigned char)(1u << j);
  }
}

void cuda_finalize(void) {
  const int th = t_th;
  if (th < 0 || th >= kMaxWorkers) return;
  ppsd_close(g_workers[th].ctx);
  g_workers[th].ctx = nullptr;
}

}  // extern "C"
