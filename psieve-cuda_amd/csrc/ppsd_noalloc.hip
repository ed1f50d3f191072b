// ppsd_noalloc.hip -- ppsd_walk_primes_noalloc: the walk and filter of
// device-resident primes into a caller's device buffer, for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It allocates nothing:
// the raw-hit ring, its counters and the filter's small-prime table live in
// `scratch`, one file-scope device array per GPU, held until the process
// ends.  Every call on a GPU shares that GPU's array, so calls that overlap
// on one GPU race; each call clears the counters it uses before its first
// kernel.
#include <mutex>

#include "ppsd_ctx.h"

using namespace ppsd;

namespace {

// Layout of the scratch array: counters, the filter's table and its powers
// of two, the raw ring.
constexpr size_t kCountersOff = 0;  // u64: raw slots requested, kept slots requested, bad primes
constexpr size_t kTableOff = 256;
constexpr size_t kTableMax = 3584;  // SmallPrime entries (3520 used: primes 11..32749 padded to 64s)
constexpr size_t kPowOff = kTableOff + kTableMax * sizeof(SmallPrime);
constexpr size_t kRingOff = kPowOff + kTableMax * kPowPer * sizeof(uint16_t);
constexpr uint64_t kRingHits = PPSD_NOALLOC_RING_HITS;
constexpr size_t kScratchBytes = kRingOff + kRingHits * sizeof(Hit);
static_assert(kPowOff % 256 == 0 && kRingOff % 256 == 0, "table and ring alignment");

}  // namespace

// The allocation-free path's device memory: one copy per GPU, at its own
// address there.  Host code never uses this name as an address: set-up
// takes each GPU's address with hipGetSymbolAddress (g_ptr below).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_ptr[kMaxDevices];  // scratch's address on each device, from hipGetSymbolAddress
uint32_t g_nsmall[kMaxDevices];

// Entries of P that are not odd in [pmin, 2^62) (counted into *bad).
__global__ __launch_bounds__(256) void count_bad_primes(const uint64_t *P, uint32_t n, uint64_t pmin,
                                                        unsigned long long *bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool b = i < n && (!(P[i] & 1) || P[i] < pmin || P[i] >= ((uint64_t)1 << 62));
  const uint64_t mask = __ballot(b);
  if (mask && __lane_id() == (uint32_t)(__ffsll((unsigned long long)mask) - 1))
    atomicAdd(bad, (unsigned long long)__popcll(mask));
}

// Set-up on the current device, once: the array's device address, and the
// filter's table copied into it.
int setup(int device, uint8_t **ptr, uint32_t *nsmall) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(PPSD_ERR_FAILURE, "GPU %d: hipGetSymbolAddress(scratch): %s", device, hipGetErrorString(e));
    const std::vector<SmallPrime> sm = small_prime_table();
    if (sm.size() > kTableMax) return set_err(PPSD_ERR_FAILURE, "small-prime table exceeds its scratch slot");
    HIP_TRY(hipMemcpy((uint8_t *)p + kTableOff, sm.data(), sm.size() * sizeof(SmallPrime), hipMemcpyHostToDevice));
    const std::vector<uint16_t> pw = pow2_table(sm);
    HIP_TRY(hipMemcpy((uint8_t *)p + kPowOff, pw.data(), pw.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    g_nsmall[device] = (uint32_t)sm.size();
    g_ptr[device] = (uint8_t *)p;
  }
  *ptr = g_ptr[device];
  *nsmall = g_nsmall[device];
  return 0;
}

}  // namespace

int ppsd_walk_primes_noalloc(int device, const ppsd_params *p, const uint64_t *P, uint32_t n, ppsd_factor *out,
                             uint64_t cap, int raw, void *stream, uint64_t *n_out) {
  if (!p || !n_out || (!P && n) || (!out && cap)) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  *n_out = 0;
  if (!walk_domain_ok(p))
    return set_err(PPSD_ERR_INVALID_PARAM, "parameters outside the walk's domain (use ppsd_derive_params)");
  if (n > PPSD_NOALLOC_MAX_PRIMES)
    return set_err(PPSD_ERR_INVALID_PARAM, "%u primes exceed the allocation-free limit of %u per call", n,
                   (unsigned)PPSD_NOALLOC_MAX_PRIMES);
  const int ndev = ppsd_device_count();
  if (device < 0 || device >= ndev || device >= kMaxDevices)
    return set_err(PPSD_ERR_NOT_IMPLEMENTED, "GPU %d not available (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  uint8_t *base = nullptr;
  uint32_t nsmall = 0;
  int rc = setup(device, &base, &nsmall);
  if (rc) return rc;
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  auto *counters = (unsigned long long *)(base + kCountersOff);
  auto *table = (const SmallPrime *)(base + kTableOff);
  auto *ring = (Hit *)(base + kRingOff);

  // the primes must be odd in [pmin, 2^62), as ppsd_submit_primes requires
  unsigned long long bad = 0;
  HIP_TRY(launch_zero_u64(counters, 3, st));
  hipLaunchKernelGGL(count_bad_primes, dim3((n + 255) / 256), dim3(256), 0, st, P, n, p->pmin, counters + 2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&bad, counters + 2, sizeof bad, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (bad) return set_err(PPSD_ERR_INVALID_PARAM, "%llu of the %u primes are not odd in [pmin, 2^62)", bad, n);

  // the library's default kernel shape and pre-test choice (ppsd_open's)
  const int r = 1, u = 16, wg = 64;
  const WalkForm form = choose_form(*p, r, u, 0, true);
  WalkArgs a{};
  a.count_d = nullptr;
  a.nmin = p->nmin;
  a.nmax = p->nmax;
  a.s = p->nstep;
  a.pre = form.pre;
  a.nwin = windows(*p);
  a.kmin = p->kmin;
  a.kmax = p->kmax;
  a.hits = ring;
  a.hit_count = counters;
  a.hit_cap = kRingHits;
  FilterArgs f{};
  f.hits = ring;
  f.hit_count = counters;
  f.hit_cap = kRingHits;
  f.kmin = p->kmin;
  f.kmax = p->kmax;
  f.nmin = p->nmin;
  f.nmax = p->nmax;
  f.kstep = p->kstep;
  f.koffset = p->koffset;
  f.small_primes = table;
  f.pow2 = (const uint16_t *)(base + kPowOff);
  f.nsmall = nsmall;
  f.kept = (Hit *)out;         // factors go straight to the caller's buffer
  f.kept_count = counters + 1;  // running over every piece: the next free slot
  f.kept_cap = cap;

  // Walk pieces of P; a piece whose raw hits overflow the ring is walked
  // again in halves, so the result never depends on the ring's size.
  // The pieces still to walk form a stack of halves, at most one per bit of
  // n plus one, so a fixed array holds it (no host allocation either).
  uint64_t total = 0;  // raw divisors reported so far (raw mode)
  uint32_t stack_first[40], stack_cnt[40];
  int depth = 0;
  stack_first[depth] = 0;
  stack_cnt[depth++] = n;
  HIP_TRY(launch_zero_u64(counters + 1, 1, st));
  while (depth > 0) {
    --depth;
    const uint32_t first = stack_first[depth], cnt = stack_cnt[depth];
    if (cnt == 0) continue;
    HIP_TRY(launch_zero_u64(counters, 1, st));
    a.P = P + first;
    a.count = cnt;
    a.cap = cnt;
    const uint32_t per_block = (uint32_t)(wg * r);
    HIP_TRY(launch_walk(a, p->sign, r, form.u, (cnt + per_block - 1) / per_block, (uint32_t)wg, st));
    unsigned long long nraw = 0;
    HIP_TRY(hipMemcpyAsync(&nraw, counters, sizeof nraw, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nraw > kRingHits) {
      if (cnt == 1)
        return set_err(PPSD_ERR_INSUFFICIENT_RESOURCE, "one prime has more than %llu hits (the allocation-free ring)",
                       (unsigned long long)kRingHits);
      stack_first[depth] = first + cnt / 2;  // the upper half after the lower
      stack_cnt[depth++] = cnt - cnt / 2;
      stack_first[depth] = first;
      stack_cnt[depth++] = cnt / 2;
      continue;
    }
    if (raw) {
      if (total < cap && nraw)
        HIP_TRY(hipMemcpyAsync(out + total, ring, std::min<uint64_t>(nraw, cap - total) * sizeof(Hit),
                               hipMemcpyDeviceToDevice, st));
      total += nraw;
    } else if (nraw) {
      HIP_TRY(launch_filter(f, kFilterGrid, st));
    }
  }
  if (!raw) HIP_TRY(hipMemcpyAsync(&total, counters + 1, sizeof total, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *n_out = total;
  return 0;
}
