// ppsd_api.hip -- host half of libppsd.so: the C ABI declared in
// include/ppsd.h.  Owns every device buffer of a context; launches the
// kernels of ppsd_kernels.hip on one non-blocking stream per context.
#include <thread>

#include "ppsd_ctx.h"

using namespace ppsd;

static thread_local std::string g_err;

namespace ppsd {

int set_err(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

uint64_t isqrt64(uint64_t x) {
  uint64_t r = (uint64_t)sqrt((double)x);
  while (r * r > x) --r;
  while ((r + 1) * (r + 1) <= x) ++r;
  return r;
}

// odd primes <= lim (host; lim is at most 46341 here)
std::vector<uint32_t> host_primes(uint32_t lim) {
  std::vector<uint8_t> comp(lim + 1, 0);
  std::vector<uint32_t> out;
  for (uint32_t i = 3; i <= lim; i += 2) {
    if (comp[i]) continue;
    out.push_back(i);
    for (uint64_t m = (uint64_t)i * i; m <= lim; m += 2 * i) comp[m] = 1;
  }
  return out;
}

// The filter's table: odd primes 11 <= q < 32768 with their reduction
// constants, padded to whole passes of 64 with copies of the last prime.
std::vector<SmallPrime> small_prime_table() {
  std::vector<SmallPrime> sm;
  const uint64_t two32 = (uint64_t)1 << 32;
  for (uint32_t q : host_primes(32767)) {
    if (q < 11) continue;
    uint32_t inv = q;  // Newton: q*q == 1 (mod 8), each step doubles the correct bits
    for (int i = 0; i < 4; ++i) inv *= 2u - q * inv;
    sm.push_back({q, (uint32_t)(two32 / q), (uint32_t)(two32 / (q - 1)), (uint32_t)(two32 % q), inv,
                  (uint32_t)((two32 - 1) / q), {0, 0}});
  }
  while (sm.size() % 64) sm.push_back(sm.back());
  return sm;
}

// Negative powers of two modulo each prime of that table, in the layout
// divides_tab reads (kPowPer u16 per prime): 2^-j mod q for j < kPowLo,
// then 2^(-kPowLo*j) mod q for j < kPowHi.
std::vector<uint16_t> pow2_table(const std::vector<SmallPrime> &sm) {
  std::vector<uint16_t> t(sm.size() * kPowPer);
  for (size_t i = 0; i < sm.size(); ++i) {
    const uint32_t q = sm[i].q;
    const uint32_t half = (q + 1) / 2;  // 2^-1 mod q
    uint16_t *row = &t[i * kPowPer];
    uint32_t x = 1;
    for (uint32_t j = 0; j < kPowLo; ++j, x = x * half % q) row[j] = (uint16_t)x;
    const uint32_t big = x;  // 2^-kPowLo mod q
    x = 1;
    for (uint32_t j = 0; j < kPowHi; ++j, x = x * big % q) row[kPowLo + j] = (uint16_t)x;
  }
  return t;
}

}  // namespace ppsd

static int lg2(uint64_t v) {
  int r = 0;
  while (v >>= 1) ++r;
  return r;
}



int ppsd_derive_params(ppsd_params *pr) {
  const uint64_t PMAX_MAX = (uint64_t)1 << 62;
  pr->status = 0;
  pr->msg[0] = 0;
  auto fail = [&](int code, const char *m) {
    pr->status = code;
    snprintf(pr->msg, sizeof pr->msg, "%s", m);
    return set_err(code, "%s", m);
  };
  if (pr->sign < -1 || pr->sign > 1) return fail(PPSD_ERR_INVALID_PARAM, "sign must be +1, -1 or 0");
  if (pr->kstep == 0) {
    pr->kstep = 2;
    pr->koffset = 1;
  }
  // main.c:916-932
  if (pr->pmin < 3) pr->pmin = 3;
  if (pr->pmax > PMAX_MAX) pr->pmax = PMAX_MAX;
  if (pr->pmin >= pr->pmax) {
    if (pr->pmax == 0 && pr->pmin < PMAX_MAX - 1000000000u)
      pr->pmax = pr->pmin + 1000000000u;
    else
      return fail(PPSD_ERR_FAILURE, "pmax must be greater than pmin");
  }
  // pps/app.c:655-712
  if (pr->kmin == 0 || pr->kmax == 0)
    return fail(PPSD_ERR_INVALID_PARAM, "Please specify an input file or all of kmin, kmax, and nmax");
  if (pr->kmin > pr->kmax) return fail(PPSD_ERR_INVALID_PARAM, "kmin <= kmax is required");
  if (pr->kmax >= pr->pmin) return fail(PPSD_ERR_INVALID_PARAM, "kmax < pmin is required");
  if (pr->kmax - pr->kmin >= ((uint64_t)3 << 36)) return fail(PPSD_ERR_INVALID_PARAM, "kmax-kmin < 3*2^36 is required");
  // -n / -N parse as unsigned at most 2^31-1 (pps/app.c:579-580)
  if (pr->nmin > 0x7fffffffu || pr->nmax > 0x7fffffffu)
    return fail(PPSD_ERR_INVALID_PARAM, "nmin and nmax must be at most 2^31-1");
  if (pr->nmin == 0) pr->nmin = (uint32_t)(2 * lg2(pr->pmin) - lg2(pr->kmax) - 1);
  if (pr->nmax == 0) pr->nmax = pr->nmin;
  if (pr->nmin > pr->nmax) return fail(PPSD_ERR_INVALID_PARAM, "nmin <= nmax is required");
  const uint64_t b0 = pr->kmin / pr->kstep, b1 = pr->kmax / pr->kstep;
  pr->kmin = b0 * pr->kstep + pr->koffset;
  pr->kmax = b1 * pr->kstep + pr->koffset;
  uint32_t s = 1;
  while ((pr->kmax << s) < pr->pmin) ++s;
  // pps/appcu.cu:155-209, 236-239
  if (s > pr->nmax - pr->nmin + 1) s = pr->nmax - pr->nmin + 1;
  if (((uint64_t)1 << (64 - s)) > pr->pmin)
    return fail(PPSD_ERR_INVALID_PARAM, "pmin is not large enough (or nmax is close to nmin)");
  if (s >= 32 && s < 48 && ((uint64_t)1 << 32) <= pr->pmin) s = 32;
  if (lg2(pr->nmin) < 6) return fail(PPSD_ERR_INVALID_PARAM, "nmin too small (must be at least 64)");
  if (pr->pmin <= ((uint64_t)1 << 32)) return fail(PPSD_ERR_INVALID_PARAM, "PMin is too small, <= 2^32");
  pr->nstep_ref = s;
  // The factor set does not depend on the window as long as
  // kmax * 2^(s-1) < pmin; the device step is written for s <= 32.
  pr->nstep = s > 32 ? 32 : s;
  return 0;
}

int ppsd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char *ppsd_last_error(void) { return g_err.c_str(); }

int ppsd_device_clock(int device, uint32_t *max_mhz) {
  int khz = 0;
  const hipError_t e = hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, device);
  if (e != hipSuccess) return set_err(PPSD_ERR_NOT_IMPLEMENTED, "GPU %d: %s", device, hipGetErrorString(e));
  if (max_mhz) *max_mhz = (uint32_t)(khz / 1000);
  return 0;
}

int ppsd_device_info(int device, char *name, uint32_t name_len, uint32_t *cus) {
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return set_err(PPSD_ERR_NOT_IMPLEMENTED, "GPU %d: %s", device, hipGetErrorString(e));
  if (name && name_len) snprintf(name, name_len, "%s", prop.gcnArchName);
  if (cus) *cus = (uint32_t)prop.multiProcessorCount;
  return 0;
}

// What the walk itself needs (ppsd_derive_params gives a subset of this):
// one candidate per window (kmax * 2^(s-1) < pmin), 2^32 < p < 2^62,
// nmin >= 64 for the 2^-nmin set-up, nmax <= 2^31-1 so that no window start
// n + U*s or hit exponent n + i wraps its 32 bits.
bool ppsd::walk_domain_ok(const ppsd_params *p) {
  return !(p->status != 0 || p->nstep < 1 || p->nstep > 32 || p->pmin <= ((uint64_t)1 << 32) ||
           p->pmax > ((uint64_t)1 << 62) || p->pmin >= p->pmax || p->nmin < 64 || p->nmin > p->nmax ||
           p->nmax > 0x7fffffffu ||
           p->kmin < 1 || p->kmin > p->kmax || (p->kmax >> (64 - p->nstep)) != 0 ||
           (p->kmax << (p->nstep - 1)) >= p->pmin || p->sign < -1 || p->sign > 1 || p->kstep < 1 ||
           p->koffset >= p->kstep);
}

static int free_ctx(ppsd_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  (void)hipFree(c->d_P);
  (void)hipFree(c->d_count);
  (void)hipFree(c->d_counters);
  (void)hipFree(c->d_hits);
  (void)hipFree(c->d_kept);
  (void)hipFree(c->d_small);
  (void)hipFree(c->d_pow2);
  (void)hipFree(c->d_spinv);
  (void)hipFree(c->d_sp);
  (void)hipFree(c->d_bitmap);
  (void)hipFree(c->d_cand);
  (void)hipFree(c->d_blockcount);
  (void)hipFree(c->d_blocksum);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (int i = 0; i < kPools; ++i) {
    if (c->b_exec[i]) (void)hipGraphExecDestroy(c->b_exec[i]);
    if (c->b_graph[i]) (void)hipGraphDestroy(c->b_graph[i]);
    (void)hipFree(c->b_pool[i]);
    (void)hipFree(c->b_pool_count[i]);
    (void)hipFree(c->b_off[i]);
    (void)hipFree(c->b_sums[i]);
    for (auto &e : c->ev2[i])
      if (e) (void)hipEventDestroy(e);
    if (c->h_off[i]) (void)hipHostFree(c->h_off[i]);
    if (c->h_sums[i]) (void)hipHostFree(c->h_sums[i]);
    (void)hipFree(c->d_rpieces[i]);
    (void)hipFree(c->d_rsegs[i]);
    (void)hipFree(c->d_rorder[i]);
    if (c->h_rorder[i]) (void)hipHostFree(c->h_rorder[i]);
    if (c->h_rpieces[i]) (void)hipHostFree(c->h_rpieces[i]);
    if (c->h_rsegs[i]) (void)hipHostFree(c->h_rsegs[i]);
  }
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  (void)hipFree(c->b_pieces);
  (void)hipFree(c->b_tasks);
  (void)hipFree(c->b_ntasks);
  (void)hipFree(c->b_dclasses);
  (void)hipFree(c->b_next);
  (void)hipFree(c->b_hits);
  (void)hipFree(c->b_kept);
  (void)hipFree(c->b_out);
  (void)hipFree(c->b_counts);
  (void)hipFree(c->b_rbitmap);
  (void)hipFree(c->b_rblockcount);
  (void)hipFree(c->b_rblocksum);
  if (c->h_pieces) (void)hipHostFree(c->h_pieces);
  if (c->h_classes) (void)hipHostFree(c->h_classes);
  if (c->h_bstage) (void)hipHostFree(c->h_bstage);
  if (c->h_out) (void)hipHostFree(c->h_out);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->ev_emit) (void)hipEventDestroy(c->ev_emit);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

constexpr uint32_t kLdsPrimeMax = (1u << 21) - 1;

// Sieving primes q <= qmax on the device: segmented sieve of [3, qmax] with
// the host's odd primes <= isqrt(qmax), compacted in order into d_sp.
static int build_sieving_primes(ppsd_ctx *c) {
  const uint64_t qmax = c->qmax;
  const uint32_t base_lim = (uint32_t)isqrt64(qmax);
  std::vector<uint32_t> base = host_primes(base_lim < 3 ? 3 : base_lim);
  // pi(x) < 1.25506 x / ln x (Rosser-Schoenfeld)
  const uint64_t cap = qmax < 17 ? 8 : (uint64_t)(1.25506 * (double)qmax / log((double)qmax)) + 16;
  uint32_t *d_base = nullptr;
  HIP_TRY(hipMalloc(&c->d_sp, cap * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&d_base, (base.size() + 1) * sizeof(uint32_t)));
  if (!base.empty())
    HIP_TRY(hipMemcpyAsync(d_base, base.data(), base.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_zero_u32(c->d_count, 1, c->stream));
  const uint64_t hi = qmax + 1;  // sieve [3, qmax]
  for (uint64_t lo = 3; lo < hi;) {
    uint64_t nbits = (hi - lo + 1) / 2;
    if (nbits > c->nbits_max) nbits = c->nbits_max;
    SieveArgs sa{};
    sa.lo = lo;
    sa.nbits = nbits;
    sa.q = d_base;
    sa.nq_small = (uint32_t)base.size();
    sa.nq = (uint32_t)base.size();
    sa.bitmap = c->d_bitmap;
    HIP_TRY(launch_sieve(sa, c->stream));
    HIP_TRY(launch_compact32(c->d_bitmap, lo, nbits, c->d_blockcount, c->d_sp, (uint32_t)cap, c->d_count,
                             c->stream));
    lo += 2 * nbits;
  }
  uint32_t n = 0;
  HIP_TRY(hipMemcpyAsync(&n, c->d_count, sizeof n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(d_base);
  if (n > cap) return set_err(PPSD_ERR_FAILURE, "sieving prime count %u exceeds bound %llu", n, (unsigned long long)cap);
  c->nsp = n;
  // Primes below 2^21 are marked per LDS segment, each segment taking one
  // remainder per prime (by a reciprocal from the host); the rest with
  // global atomics.  Per chunk of 2^27 candidates that is 512 x 1.56e5
  // remainders against the 3.6e7 atomics the primes 2^16..2^21 would cost;
  // measured on MI355X, 2^20 to 2^22 are within 5 % of each other, 2^24 is
  // 60 % slower.
  const std::vector<uint32_t> lds = host_primes((uint32_t)std::min<uint64_t>(qmax, kLdsPrimeMax));
  c->nsp_small = (uint32_t)std::min<size_t>(lds.size(), n);
  (void)hipFree(c->d_spinv);
  c->d_spinv = nullptr;
  if (c->nsp_small) {
    std::vector<uint64_t> inv(c->nsp_small);
    for (uint32_t i = 0; i < c->nsp_small; ++i) inv[i] = UINT64_MAX / lds[i];
    HIP_TRY(hipMalloc(&c->d_spinv, inv.size() * sizeof(uint64_t)));
    HIP_TRY(hipMemcpy(c->d_spinv, inv.data(), inv.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  return 0;
}

int ppsd_open(int device, const ppsd_params *p, uint32_t batch, ppsd_ctx **out) {
  if (!p || !out) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  *out = nullptr;
  if (!walk_domain_ok(p))
    return set_err(PPSD_ERR_INVALID_PARAM, "parameters outside the walk's domain (use ppsd_derive_params)");
  int ndev = ppsd_device_count();
  if (device < 0 || device >= ndev) return set_err(PPSD_ERR_NOT_IMPLEMENTED, "GPU %d not available (%d devices)", device, ndev);
  ppsd_ctx *c = new ppsd_ctx;
  c->device = device;
  c->prm = *p;
  c->batch = batch ? batch : (1u << 21);
  c->hit_cap = c->hit_alloc = 1u << 24;
  c->nbits_max = (uint64_t)1 << 27;
  int rc = 0;
  auto bail = [&](int code) {
    std::string keep = g_err;
    free_ctx(c);
    g_err = keep;
    return code;
  };
#define OPEN_TRY(expr)                                                                                \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return bail(set_err(e_ == hipErrorOutOfMemory ? PPSD_ERR_INSUFFICIENT_RESOURCE : PPSD_ERR_FAILURE, \
                          "HIP error: %s: %s", #expr, hipGetErrorString(e_)));                        \
  } while (0)
  OPEN_TRY(hipSetDevice(device));
  OPEN_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto &e : c->ev) OPEN_TRY(hipEventCreate(&e));
  OPEN_TRY(hipMalloc(&c->d_P, (size_t)c->batch * sizeof(uint64_t)));
  OPEN_TRY(hipMalloc(&c->d_count, sizeof(uint32_t)));
  OPEN_TRY(hipMalloc(&c->d_counters, 4 * sizeof(unsigned long long)));
  OPEN_TRY(hipMalloc(&c->d_hits, c->hit_cap * sizeof(Hit)));
  OPEN_TRY(hipMalloc(&c->d_kept, c->hit_cap * sizeof(Hit)));
  OPEN_TRY(hipMalloc(&c->d_bitmap, (c->nbits_max / 32 + 1) * sizeof(uint32_t)));
  OPEN_TRY(hipMalloc(&c->d_blockcount, (c->nbits_max / 32 / kScanBlock + 2) * sizeof(uint32_t)));
  OPEN_TRY(hipMalloc(&c->d_blocksum, (c->nbits_max / 32 / kScanBlock + 2) * sizeof(unsigned long long)));
  OPEN_TRY(hipHostMalloc((void **)&c->h_stage, (size_t)c->batch * sizeof(uint64_t), hipHostMallocDefault));
  {
    const std::vector<SmallPrime> sm = small_prime_table();
    c->nsmall = (uint32_t)sm.size();
    OPEN_TRY(hipMalloc(&c->d_small, sm.size() * sizeof(SmallPrime)));
    OPEN_TRY(hipMemcpy(c->d_small, sm.data(), sm.size() * sizeof(SmallPrime), hipMemcpyHostToDevice));
    const std::vector<uint16_t> pw = pow2_table(sm);
    OPEN_TRY(hipMalloc(&c->d_pow2, pw.size() * sizeof(uint16_t)));
    OPEN_TRY(hipMemcpy(c->d_pow2, pw.data(), pw.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  // Sieving primes: every odd q <= isqrt(pmax-1), or only those <= -Q when
  // given (main.c:971-973; primes up to 61 are always used, sieve.c's
  // presieve: sieve.c:770-773).
  c->qmax = isqrt64(p->pmax - 1);
  if (p->qmax) c->qmax = std::min<uint64_t>(c->qmax, std::max<uint64_t>(p->qmax, 61));
  rc = build_sieving_primes(c);
  if (rc) return bail(rc);
#undef OPEN_TRY
  *out = c;
  return 0;
}

uint32_t ppsd_batch_size(const ppsd_ctx *c) { return c ? c->batch : 0; }

int ppsd_set_tuning(ppsd_ctx *c, int r, int u, int wg) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (r && r != 1 && r != 2) return set_err(PPSD_ERR_INVALID_PARAM, "r must be 1 or 2");
  if (u && u != 4 && u != 8 && u != 16) return set_err(PPSD_ERR_INVALID_PARAM, "u must be 4, 8 or 16");
  if (wg && (wg < 64 || wg > 256 || wg % 64)) return set_err(PPSD_ERR_INVALID_PARAM, "wg must be 64..256, multiple of 64");
  if (r) c->tune_r = r;
  if (u) c->tune_u = u;
  if (wg) c->tune_wg = wg;
  return 0;
}

int ppsd_get_tuning(const ppsd_ctx *c, int *r, int *u, int *wg) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (r) *r = c->tune_r;
  if (u) *u = c->tune_u;
  if (wg) *wg = c->tune_wg;
  return 0;
}

int ppsd_set_pretest(ppsd_ctx *c, int mode) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (mode < 0 || mode > 3) return set_err(PPSD_ERR_INVALID_PARAM, "pre-test mode must be 0, 1, 2 or 3");
  c->pre_mode = mode;
  return 0;
}

int ppsd_get_pretest(const ppsd_ctx *c, int *mode, int *plain_form, int *plain_u) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  const WalkForm form = choose_form(c->prm, c->tune_r, c->tune_u, c->pre_mode, true);
  if (mode) *mode = c->pre_mode;
  if (plain_form) *plain_form = form.pre == PRE_RANGE ? 2 : form.pre == PRE_ODDLO ? 3 : 1;
  if (plain_u) *plain_u = form.u;
  return 0;
}

int ppsd_walk_form(const ppsd_params *p, int r, int u, int batched, int *form, int *chunk) {
  if (!p) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  if (!walk_domain_ok(p)) return set_err(PPSD_ERR_INVALID_PARAM, "parameters outside the walk's domain");
  if ((r != 1 && r != 2) || (u != 4 && u != 8 && u != 16))
    return set_err(PPSD_ERR_INVALID_PARAM, "r must be 1 or 2, u 4, 8 or 16");
  const WalkForm f = choose_form(*p, r, u, 0, !batched);
  if (form) *form = f.pre == PRE_RANGE ? 2 : f.pre == PRE_ODDLO ? 3 : 1;
  if (chunk) *chunk = f.u;
  return 0;
}

int ppsd_set_hit_capacity(ppsd_ctx *c, uint64_t cap) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (cap < 1 || cap > c->hit_alloc) return set_err(PPSD_ERR_INVALID_PARAM, "capacity must be 1..%llu", (unsigned long long)c->hit_alloc);
  if (c->pending != Pending::none) return set_err(PPSD_ERR_STATE, "submission outstanding");
  c->hit_cap = cap;
  return 0;
}

int ppsd_set_candidates(ppsd_ctx *c, const uint8_t *bitmap, uint32_t n0, uint32_t rows, uint64_t b0,
                        uint64_t row_bytes) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (c->pending != Pending::none) return set_err(PPSD_ERR_STATE, "a submission is outstanding");
  HIP_TRY(hipSetDevice(c->device));
  (void)hipFree(c->d_cand);
  c->d_cand = nullptr;
  c->cand_rows = 0;
  if (!bitmap) return 0;
  if (!rows || !row_bytes) return set_err(PPSD_ERR_INVALID_PARAM, "empty candidate list");
  const size_t bytes = (size_t)rows * row_bytes;
  const hipError_t e = hipMalloc(&c->d_cand, bytes);
  if (e != hipSuccess) {
    c->d_cand = nullptr;
    return set_err(PPSD_ERR_INSUFFICIENT_RESOURCE, "candidate list of %zu bytes: %s", bytes, hipGetErrorString(e));
  }
  HIP_TRY(hipMemcpy(c->d_cand, bitmap, bytes, hipMemcpyHostToDevice));
  c->cand_row_bytes = row_bytes;
  c->cand_b0 = b0;
  c->cand_n0 = n0;
  c->cand_rows = rows;
  return 0;
}

int ppsd_get_stats(const ppsd_ctx *c, ppsd_stats *st) {
  if (!c || !st) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  *st = c->st;
  return 0;
}

int ppsd_reset_stats(ppsd_ctx *c) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  c->st = ppsd_stats{};
  return 0;
}

uint32_t ppsd::windows(const ppsd_params &p) {
  return p.nmax > p.nmin ? (p.nmax - p.nmin + p.nstep - 1) / p.nstep : 0;
}

// Walk primes d_P[first .. first+n) (or the device count when count_d is set)
// into the raw ring, then filter; counters must be zero.
static int enqueue_walk(ppsd_ctx *c, uint32_t first, uint32_t n, const uint32_t *count_d, uint32_t grid_primes) {
  WalkArgs a{};
  a.P = c->d_P + first;
  a.count_d = count_d;
  a.count = n;
  a.cap = count_d ? c->batch : n;
  a.nmin = c->prm.nmin;
  a.nmax = c->prm.nmax;
  a.s = c->prm.nstep;
  const WalkForm form = choose_form(c->prm, c->tune_r, c->tune_u, c->pre_mode, true);
  a.pre = form.pre;
  a.nwin = windows(c->prm);
  a.kmin = c->prm.kmin;
  a.kmax = c->prm.kmax;
  a.hits = c->d_hits;
  a.hit_count = c->d_counters;
  a.hit_cap = c->hit_cap;
  const uint32_t per_block = (uint32_t)(c->tune_wg * c->tune_r);
  const uint32_t grid = (grid_primes + per_block - 1) / per_block;
  HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  if (grid) HIP_TRY(launch_walk(a, c->prm.sign, c->tune_r, form.u, grid, (uint32_t)c->tune_wg, c->stream));
  HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  FilterArgs f{};
  f.hits = c->d_hits;
  f.hit_count = c->d_counters;
  f.hit_cap = c->hit_cap;
  f.kmin = c->prm.kmin;
  f.kmax = c->prm.kmax;
  f.nmin = c->prm.nmin;
  f.nmax = c->prm.nmax;
  f.kstep = c->prm.kstep;
  f.koffset = c->prm.koffset;
  f.small_primes = c->d_small;
  f.pow2 = c->d_pow2;
  f.nsmall = c->nsmall;
  f.kept = c->d_kept;
  f.kept_count = c->d_counters + 1;
  f.kept_cap = c->hit_cap;
  f.cand = c->d_cand;
  f.cand_row_bytes = c->cand_row_bytes;
  f.cand_b0 = c->cand_b0;
  f.cand_n0 = c->cand_n0;
  f.cand_rows = c->cand_rows;
  HIP_TRY(hipEventRecord(c->ev[4], c->stream));
  HIP_TRY(launch_filter(f, kFilterGrid, c->stream));
  HIP_TRY(hipEventRecord(c->ev[5], c->stream));
  c->st.walk_launches++;
  return 0;
}

int ppsd_submit_primes(ppsd_ctx *c, const uint64_t *P, uint32_t n) {
  if (!c || (!P && n)) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  if (c->pending != Pending::none) return set_err(PPSD_ERR_STATE, "previous submission not collected");
  if (n > c->batch) return set_err(PPSD_ERR_INVALID_PARAM, "%u primes exceed the batch size %u", n, c->batch);
  const uint64_t top = (uint64_t)1 << 62;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t p = P[i];
    if (!(p & 1) || p < c->prm.pmin || p >= top)
      return set_err(PPSD_ERR_INVALID_PARAM, "P[%u]=%llu is not odd in [pmin, 2^62)", i, (unsigned long long)p);
  }
  HIP_TRY(hipSetDevice(c->device));
  memcpy(c->h_stage, P, (size_t)n * sizeof(uint64_t));
  HIP_TRY(hipMemcpyAsync(c->d_P, c->h_stage, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_zero_u64(c->d_counters, 4, c->stream));
  int rc = enqueue_walk(c, 0, n, nullptr, n);
  if (rc) return rc;
  c->pending = Pending::primes;
  c->n_primes_host = n;
  c->sieved = false;
  c->have_result = false;
  uint64_t sum = 0;
  for (uint32_t i = 0; i < n; ++i) sum += P[i];
  c->last_primes = n;
  c->last_sum = sum;
  return 0;
}

ppsd::WalkForm ppsd::choose_form(const ppsd_params &p, int r, int u, int mode, bool plain) {
  const bool oddlo_ok = (p.kmax >> 32) == 0;
  if (mode == 1) return {PRE_EXACT, u};
  if (mode == 2) return {PRE_RANGE, u};
  if (mode == 3) return {oddlo_ok ? PRE_ODDLO : PRE_EXACT, u};
  // false alarms per lane-window (the smallest p of the unit fires most)
  const int both = p.sign == 0 ? 1 : 0;
  const double signs = both ? 2.0 : 1.0;
  const double f_range = signs * (double)p.kmax * ldexp(1.0, (int)p.nstep - 1) / (double)p.pmin;
  const double f_odd = signs * ldexp((double)p.kmax + 1.0, -32);
  // Time per window relative to the exact form's at u = 16, measured on
  // MI355X (r = 1, p from 2^50 to 2^61), at u = 4 / 8 / 16 and negligible
  // false alarms: the range form 0.70 / 0.66 / 0.64 (one sign) and 0.455 /
  // 0.419 / 0.401 (both signs); the odd-low form 0.98 / 0.94 / 0.90 and
  // 0.94 / 0.90 / 0.86.
  static const double base_range[2][3] = {{0.70, 0.66, 0.64}, {0.455, 0.419, 0.401}};
  static const double base_odd[2][3] = {{0.98, 0.94, 0.90}, {0.94, 0.90, 0.86}};
  WalkForm best{PRE_EXACT, u};
  double best_cost = 1.0;
  for (int cu = u; cu >= 4 && (plain || cu == u); cu /= 2) {
    const int lvl = cu >= 16 ? 2 : cu >= 8 ? 1 : 0;
    // A false alarm costs, in exact window-times: a range chunk of the plain
    // walk is re-checked from its kept states, about 2.0 + 0.1 u; other
    // chunks are walked again, about u/2 with the range compare and u with
    // the exact pre-test that odd-low alarms use.  A wave's chunk holds
    // 64 r u states, so the cost per window is 64 r f times that over u.
    // The range form must win by a margin of 0.03, the odd-low form by 0.01.
    const double range_alarm = plain ? 2.0 + 0.1 * cu : 0.5 * cu;
    const double range = base_range[both][lvl] + f_range * 64.0 * r * range_alarm + 0.03;
    const double odd = base_odd[both][lvl] + f_odd * 64.0 * r * cu + 0.01;
    if (range < best_cost) {
      best_cost = range;
      best = {PRE_RANGE, cu};
    }
    if (oddlo_ok && odd < best_cost) {
      best_cost = odd;
      best = {PRE_ODDLO, cu};
    }
  }
  return best;
}

double ppsd::hits_per_prime(const ppsd_params &p, uint64_t lo) {
  const double kodd = (double)(p.kmax - p.kmin) / 2 + 1;
  return (double)windows(p) * p.nstep * kodd / (double)lo * (p.sign == 0 ? 2 : 1);
}

// Odd candidates from lo whose expected prime count (2/ln x of them),
// plus six standard deviations, fits the batch, and whose expected raw hits
// fill at most half the ring.  A rarer excess is handled in finish() by
// sieving the chunk again in halves (primes) or walking it in halves (hits).
static uint64_t range_fit(const ppsd_ctx *c, uint64_t lo) {
  const double dens = 2.0 / log((double)lo);
  const double room = (double)c->batch - 32.0;
  if (room <= 1.0) return 1;
  const double r = (-6.0 + sqrt(36.0 + 4.0 * room)) / 2.0;  // mu + 6 sqrt(mu) = room
  double fit = r * r / dens;
  const double hpp = hits_per_prime(c->prm, lo);
  if (hpp > 0) fit = std::min(fit, 0.5 * (double)c->hit_cap / hpp / dens);
  fit = std::min(fit, (double)c->nbits_max);
  return fit >= 1 ? (uint64_t)fit : 1;
}

// Sieve odd candidates [lo, lo + 2 nbits) into d_P (count in d_count).
static int enqueue_sieve(ppsd_ctx *c, uint64_t lo, uint64_t nbits) {
  HIP_TRY(launch_zero_u32(c->d_count, 1, c->stream));
  if (!nbits) return 0;
  SieveArgs sa{};
  sa.lo = lo;
  sa.nbits = nbits;
  sa.q = c->d_sp;
  sa.nq_small = c->nsp_small;
  sa.nq = c->nsp;
  sa.bitmap = c->d_bitmap;
  sa.qinv = c->d_spinv;
  HIP_TRY(hipEventRecord(c->ev[2], c->stream));
  HIP_TRY(launch_sieve(sa, c->stream));
  HIP_TRY(launch_compact64(c->d_bitmap, lo, nbits, c->d_blockcount, c->d_blocksum, c->d_P, c->batch, c->d_count,
                           c->d_counters + 2,
                           c->stream));
  HIP_TRY(hipEventRecord(c->ev[3], c->stream));
  return 0;
}

int ppsd_submit_range(ppsd_ctx *c, uint64_t p0, uint64_t p1, uint64_t *p_next) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (c->pending != Pending::none) return set_err(PPSD_ERR_STATE, "previous submission not collected");
  if (p0 < c->prm.pmin || p1 > c->prm.pmax || p0 > p1)
    return set_err(PPSD_ERR_INVALID_PARAM, "range [%llu,%llu) outside [pmin,pmax)", (unsigned long long)p0,
                   (unsigned long long)p1);
  HIP_TRY(hipSetDevice(c->device));
  const uint64_t lo = p0 | 1;
  uint64_t nbits = p1 > lo ? (p1 - lo + 1) / 2 : 0;
  nbits = std::min<uint64_t>(nbits, range_fit(c, lo));
  const uint64_t stop = nbits ? std::min<uint64_t>(lo + 2 * nbits, p1) : p1;
  HIP_TRY(launch_zero_u64(c->d_counters, 4, c->stream));
  int rc = enqueue_sieve(c, lo, nbits);
  if (rc) return rc;
  rc = enqueue_walk(c, 0, 0, c->d_count, c->batch);
  if (rc) return rc;
  c->pending = Pending::range;
  c->range_lo = lo;
  c->range_nbits = nbits;
  c->sieved = nbits != 0;
  c->have_result = false;
  if (p_next) *p_next = stop;
  return 0;
}

float ppsd::elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

// Walk d_P[first..first+n) synchronously and append its hits to the result,
// halving the piece whenever the raw ring overflows.
static int walk_pieces(ppsd_ctx *c, bool raw, uint32_t first, uint32_t n, uint64_t &raw_total, uint64_t &kept_total) {
  std::vector<std::pair<uint32_t, uint32_t>> todo{{first, n}};
  while (!todo.empty()) {
    auto [f, k] = todo.back();
    todo.pop_back();
    if (k == 0) continue;
    unsigned long long cnt[2];
    HIP_TRY(launch_zero_u64(c->d_counters, 2, c->stream));
    int rc = enqueue_walk(c, f, k, nullptr, k);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->st.walk_ms += elapsed(c->ev[0], c->ev[1]);
    c->st.filter_ms += elapsed(c->ev[4], c->ev[5]);
    if (cnt[0] > c->hit_cap) {
      if (k == 1)
        return set_err(PPSD_ERR_INSUFFICIENT_RESOURCE, "one prime has more than %llu hits",
                       (unsigned long long)c->hit_cap);
      c->st.retries++;
      todo.push_back({f + k / 2, k - k / 2});
      todo.push_back({f, k / 2});
      continue;
    }
    const uint64_t take = raw ? cnt[0] : cnt[1];
    const size_t old = c->result.size();
    c->result.resize(old + take);
    if (take)
      HIP_TRY(hipMemcpy(c->result.data() + old, raw ? c->d_hits : c->d_kept, take * sizeof(Hit), hipMemcpyDeviceToHost));
    raw_total += cnt[0];
    kept_total += cnt[1];
  }
  return 0;
}

// A range submission whose chunk held more primes than the batch: sieve and
// walk [lo, lo + 2 nbits) again in halves until each piece fits.
static int rewalk_range(ppsd_ctx *c, bool raw, uint64_t lo, uint64_t nbits, uint64_t &raw_total,
                        uint64_t &kept_total) {
  std::vector<std::pair<uint64_t, uint64_t>> todo{{lo, nbits}};
  uint64_t primes = 0, sum = 0;
  while (!todo.empty()) {
    auto [l, nb] = todo.back();
    todo.pop_back();
    if (nb == 0) continue;
    unsigned long long s2 = 0;
    uint32_t n = 0;
    HIP_TRY(launch_zero_u64(c->d_counters, 4, c->stream));
    int rc = enqueue_sieve(c, l, nb);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(&n, c->d_count, sizeof n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&s2, c->d_counters + 2, sizeof s2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->st.sieve_ms += elapsed(c->ev[2], c->ev[3]);
    if (n > c->batch) {
      todo.push_back({l + 2 * (nb / 2), nb - nb / 2});
      todo.push_back({l, nb / 2});
      continue;
    }
    rc = walk_pieces(c, raw, 0, n, raw_total, kept_total);
    if (rc) return rc;
    primes += n;
    sum += s2;
  }
  c->last_primes = primes;
  c->last_sum = sum;
  return 0;
}

// Wait for the submission; gather kept (and raw) hits, re-walking in halves
// when the raw ring overflowed.
static int finish(ppsd_ctx *c, bool raw) {
  if (c->pending == Pending::none) {
    if (c->have_result && c->result_raw == raw) return 0;
    return set_err(PPSD_ERR_STATE, "nothing submitted");
  }
  HIP_TRY(hipSetDevice(c->device));
  unsigned long long cnt[4];
  uint32_t nprimes = c->n_primes_host;
  HIP_TRY(hipMemcpyAsync(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
  if (c->pending == Pending::range) HIP_TRY(hipMemcpyAsync(&nprimes, c->d_count, sizeof nprimes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->st.walk_ms += elapsed(c->ev[0], c->ev[1]);
  c->st.filter_ms += elapsed(c->ev[4], c->ev[5]);
  if (c->sieved) c->st.sieve_ms += elapsed(c->ev[2], c->ev[3]);
  c->result.clear();
  c->result_raw = raw;
  uint64_t raw_total = 0, kept_total = 0;
  if (c->pending == Pending::range) {
    c->last_primes = nprimes;
    c->last_sum = cnt[2];
    if (nprimes > c->batch) {
      // more primes than the batch holds: sieve and walk the chunk again in
      // pieces small enough to fit, synchronously
      c->st.retries++;
      int rc = rewalk_range(c, raw, c->range_lo, c->range_nbits, raw_total, kept_total);
      c->pending = Pending::none;
      if (rc) return rc;
      c->st.primes += c->last_primes;
      c->st.prime_sum += c->last_sum;
      c->st.window_steps += c->last_primes * windows(c->prm) * (c->prm.sign == 0 ? 2 : 1);
      c->st.raw_hits += raw_total;
      c->st.factors += kept_total;
      c->have_result = true;
      return 0;
    }
  }
  if (cnt[0] <= c->hit_cap) {
    const uint64_t take = raw ? cnt[0] : cnt[1];
    c->result.resize(take);
    if (take)
      HIP_TRY(hipMemcpy(c->result.data(), raw ? c->d_hits : c->d_kept, take * sizeof(Hit), hipMemcpyDeviceToHost));
    raw_total = cnt[0];
    kept_total = cnt[1];
  } else {
    c->st.retries++;
    int rc = walk_pieces(c, raw, 0, nprimes, raw_total, kept_total);
    if (rc) {
      c->pending = Pending::none;
      return rc;
    }
  }
  const uint64_t signs = c->prm.sign == 0 ? 2 : 1;
  c->st.primes += nprimes;
  c->st.prime_sum += c->last_sum;
  c->st.window_steps += (uint64_t)nprimes * windows(c->prm) * signs;
  c->st.raw_hits += raw_total;
  c->st.factors += kept_total;
  c->pending = Pending::none;
  c->have_result = true;
  return 0;
}

static int collect(ppsd_ctx *c, bool raw, ppsd_factor *out, uint64_t cap, uint64_t *n_out, uint64_t *primes,
                   uint64_t *psum) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  int rc = finish(c, raw);
  if (rc) return rc;
  const uint64_t n = c->result.size();
  if (n_out) *n_out = n;
  if (primes) *primes = c->last_primes;
  if (psum) *psum = c->last_sum;
  static_assert(sizeof(Hit) == sizeof(ppsd_factor), "layout");
  if (out && cap) memcpy(out, c->result.data(), (size_t)std::min<uint64_t>(n, cap) * sizeof(Hit));
  return 0;
}

int ppsd_collect(ppsd_ctx *c, ppsd_factor *out, uint64_t cap, uint64_t *n_out, uint64_t *primes, uint64_t *psum) {
  return collect(c, false, out, cap, n_out, primes, psum);
}

int ppsd_collect_raw(ppsd_ctx *c, ppsd_factor *out, uint64_t cap, uint64_t *n_out, uint64_t *primes,
                     uint64_t *psum) {
  return collect(c, true, out, cap, n_out, primes, psum);
}

int ppsd_close(ppsd_ctx *c) { return free_ctx(c); }

// ---------------------------------------------------------------------------
// batched walk: many work units per launch
// ---------------------------------------------------------------------------

int ppsd::batch_alloc(ppsd_ctx *c) {
  if (c->b_ready) return 0;
  c->b_pool_cap = (uint64_t)1 << 26;
  c->b_piece_cap = 1u << 16;
  c->b_task_cap = 1u << 22;
  c->b_hit_cap = c->b_hit_alloc = (uint64_t)1 << 25;
  for (int i = 0; i < kPools; ++i) {
    HIP_TRY(hipMalloc(&c->b_pool[i], c->b_pool_cap * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&c->b_pool_count[i], sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&c->b_off[i], (c->b_piece_cap + 1) * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&c->b_sums[i], (c->b_piece_cap + 1) * sizeof(unsigned long long)));
    for (auto &e : c->ev2[i]) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipHostMalloc((void **)&c->h_off[i], (c->b_piece_cap + 1) * sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_sums[i], (c->b_piece_cap + 1) * sizeof(unsigned long long),
                          hipHostMallocDefault));
    // a round holds at most b_piece_cap pieces and fits the pool: its
    // segments number at most its bits / kSegBits + one per piece
    c->b_rseg_cap = 2 * c->b_piece_cap;
    HIP_TRY(hipMalloc(&c->d_rpieces[i], c->b_piece_cap * sizeof(RoundPiece)));
    HIP_TRY(hipMalloc(&c->d_rsegs[i], c->b_rseg_cap * sizeof(uint32_t)));
    HIP_TRY(hipHostMalloc((void **)&c->h_rpieces[i], c->b_piece_cap * sizeof(RoundPiece), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_rsegs[i], c->b_rseg_cap * sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipMalloc(&c->d_rorder[i], c->b_piece_cap * sizeof(uint32_t)));
    HIP_TRY(hipHostMalloc((void **)&c->h_rorder[i], c->b_piece_cap * sizeof(uint32_t), hipHostMallocDefault));
  }
  HIP_TRY(hipMalloc(&c->b_rbitmap, (size_t)kRoundSegs * kSegWords * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&c->b_rblockcount, (size_t)kRoundSegs * kBlocksPerSeg * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&c->b_rblocksum, (size_t)kRoundSegs * kBlocksPerSeg * sizeof(unsigned long long)));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIP_TRY(hipMalloc(&c->b_pieces, c->b_piece_cap * sizeof(PieceDesc)));
  HIP_TRY(hipMalloc(&c->b_tasks, (size_t)c->b_task_cap * sizeof(Task)));
  HIP_TRY(hipMalloc(&c->b_ntasks, sizeof(uint32_t)));
  c->b_class_cap = 4 * c->b_piece_cap;  // at most four classes a piece
  HIP_TRY(hipMalloc(&c->b_dclasses, (size_t)c->b_class_cap * sizeof(TaskClass)));
  HIP_TRY(hipHostMalloc((void **)&c->h_classes, (size_t)c->b_class_cap * sizeof(TaskClass), hipHostMallocDefault));
  HIP_TRY(hipMalloc(&c->b_next, sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&c->b_hits, c->b_hit_alloc * sizeof(BHit)));
  HIP_TRY(hipMalloc(&c->b_kept, c->b_hit_alloc * sizeof(BHit)));
  HIP_TRY(hipMalloc(&c->b_counts, 2 * sizeof(unsigned long long)));
  // results stay on the device until ppsd_batch_results: room for 2^26
  // (2 GiB), flushed to the host mid-batch only beyond that
  c->b_out_cap = (uint64_t)1 << 26;
  HIP_TRY(hipMalloc(&c->b_out, c->b_out_cap * sizeof(BHit)));
  c->h_bstage_cap = (uint64_t)1 << 20;
  HIP_TRY(hipHostMalloc((void **)&c->h_bstage, c->h_bstage_cap * sizeof(BHit), hipHostMallocDefault));
  c->h_out_cap = (uint64_t)1 << 24;  // 512 MiB
  HIP_TRY(hipHostMalloc((void **)&c->h_out, c->h_out_cap * sizeof(BHit), hipHostMallocDefault));
  HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&c->ev_emit, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc((void **)&c->h_pieces, c->b_piece_cap * sizeof(PieceDesc), hipHostMallocDefault));
  c->b_ready = true;
  return 0;
}

// Sieving primes up to isqrt(pmax-1) for the largest unit of a batch, and a
// host sample of every kSpSample-th of them to bound each piece's share.
constexpr uint32_t kSpSample = 4096;

int ppsd::ensure_sieving(ppsd_ctx *c, uint64_t pmax) {
  const uint64_t need = isqrt64(pmax - 1);
  if (need > c->qmax) {
    (void)hipFree(c->d_sp);
    c->d_sp = nullptr;
    c->qmax = need;
    c->sp_sample.clear();
    int rc = build_sieving_primes(c);
    if (rc) return rc;
  }
  const uint32_t ns = (c->nsp + kSpSample - 1) / kSpSample;
  if (c->sp_sample.size() != ns) {
    c->sp_sample.resize(ns);
    if (ns)
      HIP_TRY(hipMemcpy2D(c->sp_sample.data(), sizeof(uint32_t), c->d_sp, kSpSample * sizeof(uint32_t),
                          sizeof(uint32_t), ns, hipMemcpyDeviceToHost));
  }
  return 0;
}

// Sieving primes a chunk below hi needs (q <= isqrt(hi-1)), rounded up to
// the sample grid; any prime q < hi's chunk start is harmless to sieve with.
static uint32_t sieve_count(const ppsd_ctx *c, uint64_t hi) {
  const uint64_t lim = isqrt64(hi - 1);
  const auto it = std::upper_bound(c->sp_sample.begin(), c->sp_sample.end(), (uint32_t)std::min<uint64_t>(lim, ~0u));
  const uint64_t n = (uint64_t)(it - c->sp_sample.begin()) * kSpSample;
  return (uint32_t)std::min<uint64_t>(n, c->nsp);
}

static BatchArgs batch_args(ppsd_ctx *c, int buf) {
  BatchArgs a{};
  a.pool = c->b_pool[buf];
  a.pieces = c->b_pieces;
  a.tasks = c->b_tasks;
  a.ntasks = c->b_ntasks;
  a.next_task = c->b_next;
  a.hits = c->b_hits;
  a.hit_count = c->b_counts;
  a.hit_cap = c->b_hit_cap;
  a.kept = c->b_kept;
  a.kept_count = c->b_counts + 1;
  a.kept_cap = c->b_hit_cap;
  a.small_primes = c->d_small;
  a.pow2 = c->d_pow2;
  a.nsmall = c->nsmall;
  return a;
}

static uint32_t device_cus(ppsd_ctx *c) {
  if (!c->cus) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
      cus = 256;
    c->cus = (uint32_t)cus;
  }
  return c->cus;
}

static uint32_t persist_grid(ppsd_ctx *c) {
  return device_cus(c) * (uint32_t)c->b_waves;  // 256-thread groups: b_waves waves per SIMD
}

// The round's device sequence -- zero the counters and the queue head, drain
// the queue, filter -- as one graph per pool, captured once per kernel shape.
static int round_graph(ppsd_ctx *c, int buf) {
  const int key = c->b_r * 1000000 + c->b_u * 1000 + c->b_waves;
  if (c->b_exec[buf] && c->b_graph_key[buf] == key) return 0;
  if (c->b_exec[buf]) (void)hipGraphExecDestroy(c->b_exec[buf]);
  if (c->b_graph[buf]) (void)hipGraphDestroy(c->b_graph[buf]);
  c->b_exec[buf] = nullptr;
  c->b_graph[buf] = nullptr;
  const BatchArgs a = batch_args(c, buf);
  const uint32_t grid = persist_grid(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_zero_u64(c->b_counts, 2, c->stream);
  if (e == hipSuccess) e = launch_zero_u32(c->b_next, 1, c->stream);
  if (e == hipSuccess) e = launch_persist(a, c->b_r, c->b_u, grid, c->stream);
  if (e == hipSuccess) e = launch_batch_filter(a, kFilterGrid, c->stream);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return set_err(PPSD_ERR_FAILURE, "HIP error capturing the round: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return set_err(PPSD_ERR_FAILURE, "HIP error ending capture: %s", hipGetErrorString(e2));
  c->b_graph[buf] = g;
  HIP_TRY(hipGraphInstantiate(&c->b_exec[buf], g, nullptr, nullptr, 0));
  c->b_graph_key[buf] = key;
  return 0;
}

// Task classes of one round over the sieved pieces (c->b_classes): the
// segment length halves from b_seg_max down to b_seg_min until the queue
// holds enough tasks to keep every wave slot busy through the tail.  A
// piece's tasks come in at most four sizes -- full or last block of primes
// times full or last segment of windows -- so the round's order, largest
// task first, is its classes sorted by task size; the tasks themselves are
// written on the device (launch_expand), which spares the host millions of
// task records a round.  Returns the number of tasks.
static uint64_t build_classes(ppsd_ctx *c, uint32_t npieces) {
  const uint32_t per = 64u * (uint32_t)c->b_r;
  const uint64_t target = (uint64_t)persist_grid(c) * 4 * 8;  // waves resident x 8
  uint32_t seg = c->b_seg_max;
  for (;;) {
    uint64_t n = 0;
    for (uint32_t i = 0; i < npieces; ++i) {
      const PieceDesc &P = c->h_pieces[i];
      n += (uint64_t)((P.nprimes + per - 1) / per) * ((P.nwin + seg - 1) / seg);
    }
    if (n >= target || seg <= c->b_seg_min) break;
    seg /= 2;
  }
  std::vector<TaskClass> &cls = c->b_classes;
  cls.clear();
  for (uint32_t i = 0; i < npieces; ++i) {
    const PieceDesc &P = c->h_pieces[i];
    if (!P.nprimes || !P.nwin) continue;
    const uint32_t nfull = P.nprimes / per, flast = P.nprimes % per;  // full blocks, last block's count
    const uint32_t wfull = P.nwin / seg, wlast = P.nwin % seg;
    // nf blocks of `count` primes from f0, nseg segments of nw windows from w0
    auto add = [&](uint32_t f0, uint32_t nf, uint32_t count, uint32_t w0, uint32_t nseg, uint32_t nw) {
      if (nf && nseg && count && nw) cls.push_back(TaskClass{0, i, f0, per, nf, count, w0, nw, nseg});
    };
    add(0, nfull, per, 0, wfull, seg);
    add(0, nfull, per, wfull * seg, 1, wlast);
    add(nfull * per, 1, flast, 0, wfull, seg);
    add(nfull * per, 1, flast, wfull * seg, 1, wlast);
  }
  std::stable_sort(cls.begin(), cls.end(), [](const TaskClass &x, const TaskClass &y) {
    return (uint64_t)x.count * x.nw > (uint64_t)y.count * y.nw;
  });
  uint64_t t = 0;
  for (TaskClass &k : cls) {
    k.t_begin = t;
    t += (uint64_t)k.nf * k.nseg;
  }
  return t;
}

// A launch drains its queue largest task first, so its end is set by the
// last tasks the waves take: with equal tasks of a whole segment each, waves
// finish up to a task's length apart (a segment of 65536 windows takes a
// wave sharing its SIMD with five others several ms).  A launch's last 2
// tasks per resident wave are therefore cut into kTailParts shorter
// segments (none below kTailMinWindows windows, whose start states cost
// about 1 % of their walk) and queued after the rest, one part of each in
// turn: the waves then run out of work within one short task of each other.
constexpr uint32_t kTailParts = 8;
constexpr uint32_t kTailMinWindows = 4096;

// Copies the batch's device output so far (b_out, mirrored in h_out) to
// b_result and empties b_out.
int ppsd::batch_flush(ppsd_ctx *c) {
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamSynchronize(c->copy_stream));
  if (c->b_out_n) {
    const size_t old = c->b_result.size();
    const uint64_t mirrored = std::min(c->b_out_n, c->h_out_cap);
    c->b_result.resize(old + c->b_out_n);
    memcpy(c->b_result.data() + old, c->h_out, mirrored * sizeof(BHit));
    if (c->b_out_n > mirrored)
      HIP_TRY(hipMemcpy(c->b_result.data() + old + mirrored, c->b_out + mirrored,
                        (c->b_out_n - mirrored) * sizeof(BHit), hipMemcpyDeviceToHost));
    c->b_out_n = 0;
  }
  return 0;
}

// Run tasks [t0, t1) of the round's order (c->b_classes, on the device in
// b_dclasses) through the round graph, queued by launch_expand; re-run
// halves when the raw ring overflows.  Appends kept (or raw) hits to b_out.
static int run_tasks(ppsd_ctx *c, uint64_t t0, uint64_t t1, bool raw, int buf) {
  const uint64_t waves = (uint64_t)persist_grid(c) * 4;
  std::vector<std::pair<uint64_t, uint64_t>> todo{{t0, t1}};
  while (!todo.empty()) {
    auto [f, e] = todo.back();
    todo.pop_back();
    const uint64_t k = e - f;
    if (k == 0) continue;
    int rc = round_graph(c, buf);
    if (rc) return rc;
    ExpandArgs x{};
    x.classes = c->b_dclasses;
    x.nclasses = (uint32_t)c->b_classes.size();
    x.t0 = f;
    x.t1 = e;
    x.tail = (uint32_t)std::min<uint64_t>(k, 2 * waves);
    x.parts = kTailParts;
    x.tail_min_windows = kTailMinWindows;
    x.out = c->b_tasks;
    x.nout = c->b_ntasks;
    if ((k - x.tail) + (uint64_t)x.tail * x.parts > c->b_task_cap)
      return set_err(PPSD_ERR_FAILURE, "launch of %llu tasks exceeds the queue", (unsigned long long)k);
    HIP_TRY(launch_expand(x, c->stream));
    unsigned long long cnt[2];
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    HIP_TRY(hipGraphLaunch(c->b_exec[buf], c->stream));
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    HIP_TRY(hipMemcpyAsync(cnt, c->b_counts, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->st.walk_ms += elapsed(c->ev[0], c->ev[1]);
    c->st.walk_launches++;
    if (cnt[0] > c->b_hit_cap) {
      c->st.retries++;
      if (k == 1) {
        // one task (at most 64*R primes) needs more room than the ring
        return set_err(PPSD_ERR_INSUFFICIENT_RESOURCE, "one task has %llu hits (ring %llu)", cnt[0],
                       (unsigned long long)c->b_hit_cap);
      }
      todo.push_back({f + k / 2, e});
      todo.push_back({f, f + k / 2});
      continue;
    }
    // the results stay on the device until the batch ends or its output
    // fills; the output is used up to twice the ring (all of it at the
    // default ring), so that a small ring also exercises the flush
    const uint64_t take = raw ? cnt[0] : cnt[1];
    if (c->b_out_n + take > std::min(c->b_out_cap, 2 * c->b_hit_cap)) {
      int frc = batch_flush(c);
      if (frc) return frc;
    }
    HIP_TRY(launch_emit(raw ? c->b_hits : c->b_kept, take, c->b_pieces, c->b_out + c->b_out_n, c->stream));
    if (take && c->b_out_n < c->h_out_cap) {  // to the host mirror beside the next walk
      const uint64_t m = std::min(take, c->h_out_cap - c->b_out_n);
      HIP_TRY(hipEventRecord(c->ev_emit, c->stream));
      HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->ev_emit, 0));
      HIP_TRY(hipMemcpyAsync(c->h_out + c->b_out_n, c->b_out + c->b_out_n, m * sizeof(BHit),
                             hipMemcpyDeviceToHost, c->copy_stream));
    }
    c->b_out_n += take;
    c->st.raw_hits += cnt[0];
    c->st.factors += cnt[1];
  }
  return 0;
}

// Sieve pieces [b, e) back to back into pool `buf` on stream st
// (asynchronous): the pieces' segments go to one bitmap, sieved, compacted
// and totalled per piece in one set of launches for as many pieces as its
// kRoundSegs segments hold -- a handful of launches per round, where a chain
// of small launches per piece had left the device mostly idle between walks.
// b_off[buf][i] and b_sums[buf][i] receive each piece's pool offset and prime
// sum, b_off[buf][np] the pool count.
int ppsd::sieve_round(ppsd_ctx *c, const std::vector<Piece> &pieces, size_t b, size_t e, int buf, hipStream_t st,
                      hipEvent_t ev0, hipEvent_t ev1) {
  const uint32_t np = (uint32_t)(e - b);
  if (np > c->b_piece_cap) return set_err(PPSD_ERR_FAILURE, "%u pieces in one round (at most %u)", np, c->b_piece_cap);
  RoundPiece *rp = c->h_rpieces[buf];
  uint32_t *segs = c->h_rsegs[buf];
  // launches: [piece range, segment range) each within kRoundSegs segments
  struct Launch {
    uint32_t p0, p1, s0, s1;
    uint32_t qhi;  // the largest nq of its pieces
  };
  std::vector<Launch> launches;
  Launch cur{0, 0, 0, 0, 0};
  uint32_t nseg_all = 0;
  for (uint32_t i = 0; i < np; ++i) {
    const Piece &P = pieces[b + i];
    const uint64_t lo = P.lo | 1;
    const uint64_t nbits = P.hi > lo ? (P.hi - lo + 1) / 2 : 0;
    const uint64_t ns = (nbits + kSegBits - 1) / kSegBits;
    if (ns > kRoundSegs) return set_err(PPSD_ERR_FAILURE, "one piece needs %llu sieve segments", (unsigned long long)ns);
    if (nseg_all + ns > c->b_rseg_cap) return set_err(PPSD_ERR_FAILURE, "round needs more than %u sieve segments", c->b_rseg_cap);
    if ((nseg_all - cur.s0) + ns > kRoundSegs) {  // this piece starts the next launch
      cur.p1 = i;
      cur.s1 = nseg_all;
      launches.push_back(cur);
      cur = Launch{i, i, nseg_all, nseg_all, 0};
    }
    RoundPiece &R = rp[i];
    R.lo = lo;
    R.nbits = nbits;
    R.bit0 = (uint64_t)(nseg_all - cur.s0) * kSegBits;
    R.nq = nbits ? sieve_count(c, lo + 2 * nbits) : 0;
    R.nq_small = std::min(c->nsp_small, R.nq);
    cur.qhi = std::max(cur.qhi, R.nq);
    for (uint64_t k = 0; k < ns; ++k) segs[nseg_all++] = i;
  }
  cur.p1 = np;
  cur.s1 = nseg_all;
  launches.push_back(cur);
  // per launch, its pieces by descending nq: a large prime's thread stops at
  // the first piece that does not sieve with it
  uint32_t *order = c->h_rorder[buf];
  for (const Launch &L : launches) {
    for (uint32_t i = L.p0; i < L.p1; ++i) order[i] = i;
    std::stable_sort(order + L.p0, order + L.p1, [&](uint32_t x, uint32_t y) { return rp[x].nq > rp[y].nq; });
  }
  HIP_TRY(hipEventRecord(ev0, st));
  HIP_TRY(launch_zero_u32(c->b_pool_count[buf], 1, st));
  if (np) {
    HIP_TRY(hipMemcpyAsync(c->d_rpieces[buf], rp, np * sizeof(RoundPiece), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_rorder[buf], order, np * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (nseg_all)
      HIP_TRY(hipMemcpyAsync(c->d_rsegs[buf], segs, nseg_all * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  }
  for (const Launch &L : launches) {
    if (L.p1 == L.p0) continue;
    RoundSieveArgs a{};
    a.pieces = c->d_rpieces[buf];
    a.p0 = L.p0;
    a.p1 = L.p1;
    a.seg_piece = c->d_rsegs[buf] + L.s0;
    a.nseg = L.s1 - L.s0;
    a.by_nq = c->d_rorder[buf];
    a.qlo = c->nsp_small;  // every piece with large primes marks q[0..nsp_small) in LDS
    a.qhi = std::max(L.qhi, c->nsp_small);
    a.q = c->d_sp;
    a.qinv = c->d_spinv;
    a.bitmap = c->b_rbitmap;
    a.block = 256;  // one wave per SIMD: fits beside the persistent walk's waves
    a.grid = 2 * device_cus(c);  // two workgroups a CU, as many as the slots the walk leaves free at its default shape
    HIP_TRY(launch_round_sieve(a, c->b_rblockcount, c->b_rblocksum, c->b_pool[buf], (uint32_t)c->b_pool_cap,
                               c->b_pool_count[buf], c->b_off[buf], c->b_sums[buf], st));
  }
  HIP_TRY(hipMemcpyAsync(c->b_off[buf] + np, c->b_pool_count[buf], sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->h_off[buf], c->b_off[buf], (np + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  if (np)
    HIP_TRY(hipMemcpyAsync(c->h_sums[buf], c->b_sums[buf], np * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           st));
  HIP_TRY(hipEventRecord(ev1, st));
  return 0;
}

// Wait for the sieve of pieces [b, e) in pool `buf`, then build and drain the
// round's task queue on c->stream.  Returns 1 (no error, nothing recorded)
// when the pieces overflowed the pool and the caller must split them.
int ppsd::walk_round(ppsd_ctx *c, const std::vector<Piece> &pieces, size_t b, size_t e,
                     const std::vector<const ppsd_params *> &par, bool raw, int buf, hipEvent_t ev0,
                     hipEvent_t ev1) {
  const uint32_t np = (uint32_t)(e - b);
  HIP_TRY(hipEventSynchronize(ev1));
  c->st.sieve_ms += elapsed(ev0, ev1);
  const uint32_t *off = c->h_off[buf];
  const unsigned long long *sums = c->h_sums[buf];
  if (off[np] > c->b_pool_cap) return 1;
  for (uint32_t i = 0; i < np; ++i) {
    const Piece &P = pieces[b + i];
    const ppsd_params &u = *par[P.unit];
    PieceDesc &D = c->h_pieces[i];
    D.kmin = u.kmin;
    D.kmax = u.kmax;
    D.nmin = u.nmin;
    D.nmax = u.nmax;
    D.s = u.nstep;
    D.nwin = windows(u);
    D.sign = u.sign;
    D.kstep = u.kstep;
    D.koffset = u.koffset;
    D.prime_off = off[i];
    D.nprimes = off[i + 1] - off[i];
    D.unit = P.unit;
    D.pre = choose_form(u, c->b_r, c->b_u, c->pre_mode, false).pre;
    c->b_unit_primes[P.unit] += D.nprimes;
    c->b_unit_sums[P.unit] += sums[i];
    c->st.primes += D.nprimes;
    c->st.prime_sum += sums[i];
    c->st.window_steps += (uint64_t)D.nprimes * D.nwin * (D.sign == 0 ? 2 : 1);
  }
  HIP_TRY(hipMemcpyAsync(c->b_pieces, c->h_pieces, np * sizeof(PieceDesc), hipMemcpyHostToDevice, c->stream));
  const uint64_t ntask = build_classes(c, np);
  if (!ntask) return 0;
  const std::vector<TaskClass> &cls = c->b_classes;
  if (cls.size() > c->b_class_cap)
    return set_err(PPSD_ERR_FAILURE, "%zu task classes in one round (at most %u)", cls.size(), c->b_class_cap);
  memcpy(c->h_classes, cls.data(), cls.size() * sizeof(TaskClass));
  HIP_TRY(hipMemcpyAsync(c->b_dclasses, c->h_classes, cls.size() * sizeof(TaskClass), hipMemcpyHostToDevice,
                         c->stream));
  // launches whose expected raw hits fill at most half the ring: every task
  // of a class expects the same number, so each boundary falls at a task
  // index computed per class
  const double room = 0.5 * (double)c->b_hit_cap;
  const uint64_t waves = (uint64_t)persist_grid(c) * 4;
  const uint64_t group_max = c->b_task_cap - 2 * waves * kTailParts;  // room for the cut tail
  uint64_t g0 = 0, gn = 0;  // the launch being formed: its first task and its count
  double eh = 0;  // its expected raw hits
  for (const TaskClass &k : cls) {
    const PieceDesc &D = c->h_pieces[k.piece];
    const Piece &P = pieces[b + k.piece];
    const double hpw = D.nwin ? hits_per_prime(*par[P.unit], P.lo | 1) / D.nwin : 0.0;
    const double et = hpw * k.count * k.nw;  // expected raw hits of one task of the class
    uint64_t left = (uint64_t)k.nf * k.nseg;
    while (left) {
      uint64_t take = std::min(left, group_max - gn);
      if (et > 0) {
        const double fit = (room - eh) / et;  // tasks that still fit (may be huge)
        if (fit < (double)take) take = fit > 0 ? (uint64_t)fit : 0;
      }
      if (take == 0 && gn == 0) take = 1;
      if (take == 0) {  // the launch is full
        int rc = run_tasks(c, g0, g0 + gn, raw, buf);
        if (rc) return rc;
        g0 += gn;
        gn = 0;
        eh = 0;
        continue;
      }
      gn += take;
      eh += et * (double)take;
      left -= take;
    }
  }
  return run_tasks(c, g0, g0 + gn, raw, buf);
}

int ppsd_batch_prepare(ppsd_ctx *c, uint64_t pmax_all) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (pmax_all < 3 || pmax_all > ((uint64_t)1 << 62)) return set_err(PPSD_ERR_INVALID_PARAM, "pmax must be 3..2^62");
  if (c->prm.qmax && c->qmax < isqrt64(c->prm.pmax - 1))
    return set_err(PPSD_ERR_STATE, "context opened with a partial sieve (qmax)");
  HIP_TRY(hipSetDevice(c->device));
  int rc = batch_alloc(c);
  if (rc) return rc;
  // both pools' round graphs (captured again only when the kernel shape or
  // the ring changed), not at the first round that uses each
  for (int buf = 0; buf < kPools; ++buf) {
    rc = round_graph(c, buf);
    if (rc) return rc;
  }
  return ensure_sieving(c, pmax_all);
}

int ppsd_batch_run(ppsd_ctx *c, const ppsd_params *units, uint32_t nunits, int raw) {
  if (!c || (!units && nunits)) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  if (c->pending != Pending::none) return set_err(PPSD_ERR_STATE, "a submission is outstanding");
  std::vector<const ppsd_params *> par(nunits);
  uint64_t pmax_all = 0;
  for (uint32_t u = 0; u < nunits; ++u) {
    if (!walk_domain_ok(&units[u]))
      return set_err(PPSD_ERR_INVALID_PARAM, "unit %u: parameters outside the walk's domain", u);
    if (units[u].qmax && units[u].qmax < isqrt64(units[u].pmax - 1))
      return set_err(PPSD_ERR_INVALID_PARAM, "unit %u: a batch sieves completely (qmax must be 0)", u);
    par[u] = &units[u];
    pmax_all = std::max(pmax_all, units[u].pmax);
  }
  if (c->prm.qmax && c->qmax < isqrt64(c->prm.pmax - 1))
    return set_err(PPSD_ERR_STATE, "context opened with a partial sieve (qmax)");
  HIP_TRY(hipSetDevice(c->device));
  int rc = batch_alloc(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream2));  // nothing of an earlier, failed run still in flight
  HIP_TRY(hipStreamSynchronize(c->copy_stream));
  if (nunits) {
    rc = ppsd_batch_prepare(c, pmax_all);
    if (rc) return rc;
  }
  c->b_result.clear();
  c->b_out_n = 0;
  c->b_unit_primes.assign(nunits, 0);
  c->b_unit_sums.assign(nunits, 0);
  c->b_nunits = nunits;
  // pieces of at most a quarter of the pool (expected primes)
  std::vector<Piece> pieces;
  const double piece_max = (double)c->b_pool_cap / 4;
  for (uint32_t u = 0; u < nunits; ++u) {
    const uint64_t lo = units[u].pmin, hi = units[u].pmax;
    const double dens = 1.05 / log((double)lo);
    const uint64_t width = std::max<uint64_t>(2, (uint64_t)(piece_max / dens) & ~(uint64_t)1);
    for (uint64_t a = lo; a < hi; a += width) {
      const uint64_t z = std::min(hi, a + width);
      pieces.push_back(Piece{u, a, z, (double)(z - a) * dens + 64});
    }
  }
  // rounds: pieces packed in order up to 90% of the pool
  std::vector<std::pair<size_t, size_t>> rounds;
  for (size_t b = 0; b < pieces.size();) {
    size_t e = b;
    double est = 0;
    while (e < pieces.size() && e - b < c->b_piece_cap && (e == b || est + pieces[e].est <= 0.9 * c->b_pool_cap))
      est += pieces[e++].est;
    rounds.push_back({b, e});
    b = e;
  }
  // Round i is walked on the context's stream from pool i mod kPools while
  // rounds i+1 and i+2 are sieved on stream2 into the others; round i+2's
  // sieve is enqueued only after round i-1's walk (which read that pool)
  // has finished.
  auto sieve = [&](size_t i) {
    const int buf = (int)(i % kPools);
    return sieve_round(c, pieces, rounds[i].first, rounds[i].second, buf, c->stream2, c->ev2[buf][0],
                       c->ev2[buf][1]);
  };
  for (size_t i = 0; i < std::min<size_t>(kPools - 1, rounds.size()); ++i) {
    rc = sieve(i);
    if (rc) return rc;
  }
  for (size_t i = 0; i < rounds.size(); ++i) {
    const int buf = (int)(i % kPools);
    const auto [b, e] = rounds[i];
    if (i + kPools - 1 < rounds.size()) {
      rc = sieve(i + kPools - 1);
      if (rc) return rc;
    }
    rc = walk_round(c, pieces, b, e, par, raw != 0, buf, c->ev2[buf][0], c->ev2[buf][1]);
    if (rc == 1) {  // pool overflow (rare): this round's pieces again, in two halves, on pool `buf`
      if (e - b == 1) return set_err(PPSD_ERR_FAILURE, "one piece overflowed the prime pool");
      HIP_TRY(hipStreamSynchronize(c->stream2));  // the sieve buffers are shared with the next rounds' sieves
      const size_t mid = b + (e - b) / 2;
      for (const auto &[x, y] : {std::make_pair(b, mid), std::make_pair(mid, e)}) {
        rc = sieve_round(c, pieces, x, y, buf, c->stream, c->ev[2], c->ev[3]);
        if (!rc) rc = walk_round(c, pieces, x, y, par, raw != 0, buf, c->ev[2], c->ev[3]);
        if (rc) return rc < 0 ? rc : set_err(PPSD_ERR_FAILURE, "prime pool overflow");
      }
    } else if (rc) {
      return rc;
    }
  }
  // the results stay where they are (b_result, then b_out) until
  // ppsd_batch_results reads them
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

// Records [0, n) of src into out / unit_of, split over a few host threads
// (tens of millions of records for a hit-dense batch).
static void put_results(const BHit *src, uint64_t n, ppsd_factor *out, uint32_t *unit_of) {
  auto part = [=](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const BHit &h = src[i];
      if (out) out[i] = ppsd_factor{h.p, h.k, h.n, h.c};
      if (unit_of) unit_of[i] = h.piece;
    }
  };
  const uint64_t per_thread = (uint64_t)1 << 20;
  const unsigned nt = (unsigned)std::min<uint64_t>(8, (n + per_thread - 1) / per_thread);
  if (nt <= 1) {
    part(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) th.emplace_back(part, n * t / nt, n * (t + 1) / nt);
  for (auto &x : th) x.join();
}

int ppsd_batch_results(ppsd_ctx *c, ppsd_factor *out, uint32_t *unit_of, uint64_t cap, uint64_t *n_out) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  const uint64_t nhost = c->b_result.size();
  if (n_out) *n_out = nhost + c->b_out_n;
  const uint64_t n = std::min<uint64_t>(cap, nhost + c->b_out_n);
  if (!out && !unit_of) return 0;
  put_results(c->b_result.data(), std::min(n, nhost), out, unit_of);
  if (n <= nhost) return 0;
  // the device part: its mirrored head from pinned memory, the rest through
  // the pinned staging buffer
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->copy_stream));
  const uint64_t mirrored = std::min(n - nhost, c->h_out_cap);
  put_results(c->h_out, mirrored, out ? out + nhost : nullptr, unit_of ? unit_of + nhost : nullptr);
  for (uint64_t i = nhost + mirrored; i < n;) {
    const uint64_t m = std::min(n - i, c->h_bstage_cap);
    HIP_TRY(hipMemcpy(c->h_bstage, c->b_out + (i - nhost), m * sizeof(BHit), hipMemcpyDeviceToHost));
    put_results(c->h_bstage, m, out ? out + i : nullptr, unit_of ? unit_of + i : nullptr);
    i += m;
  }
  return 0;
}

int ppsd_batch_unit_primes(const ppsd_ctx *c, uint64_t *count, uint64_t *sum, uint32_t nunits) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (nunits > c->b_nunits) return set_err(PPSD_ERR_INVALID_PARAM, "only %u units in the last batch", c->b_nunits);
  for (uint32_t u = 0; u < nunits; ++u) {
    if (count) count[u] = c->b_unit_primes[u];
    if (sum) sum[u] = c->b_unit_sums[u];
  }
  return 0;
}

int ppsd_batch_set_tuning(ppsd_ctx *c, int r, int u, int waves_per_simd, uint32_t seg_min, uint32_t seg_max) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (r && r != 1 && r != 2) return set_err(PPSD_ERR_INVALID_PARAM, "r must be 1 or 2");
  if (u && u != 4 && u != 8 && u != 16) return set_err(PPSD_ERR_INVALID_PARAM, "u must be 4, 8 or 16");
  if (waves_per_simd && (waves_per_simd < 1 || waves_per_simd > 16))
    return set_err(PPSD_ERR_INVALID_PARAM, "waves per SIMD must be 1..16");
  if ((seg_min && seg_min < 16) || (seg_max && seg_min && seg_max < seg_min))
    return set_err(PPSD_ERR_INVALID_PARAM, "segment lengths: 16 <= seg_min <= seg_max");
  if (r) c->b_r = r;
  if (u) c->b_u = u;
  if (waves_per_simd) c->b_waves = waves_per_simd;
  if (seg_min) c->b_seg_min = seg_min;
  if (seg_max) c->b_seg_max = seg_max;
  if (c->b_seg_max < c->b_seg_min) c->b_seg_max = c->b_seg_min;
  return 0;
}

int ppsd_batch_get_tuning(const ppsd_ctx *c, int *r, int *u, int *waves_per_simd, uint32_t *seg_min,
                          uint32_t *seg_max) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  if (r) *r = c->b_r;
  if (u) *u = c->b_u;
  if (waves_per_simd) *waves_per_simd = c->b_waves;
  if (seg_min) *seg_min = c->b_seg_min;
  if (seg_max) *seg_max = c->b_seg_max;
  return 0;
}

int ppsd_batch_set_hit_capacity(ppsd_ctx *c, uint64_t cap) {
  if (!c) return set_err(PPSD_ERR_INVALID_PARAM, "null context");
  int rc = batch_alloc(c);
  if (rc) return rc;
  if (cap < 1 || cap > c->b_hit_alloc)
    return set_err(PPSD_ERR_INVALID_PARAM, "capacity must be 1..%llu", (unsigned long long)c->b_hit_alloc);
  c->b_hit_cap = cap;
  c->b_graph_key[0] = c->b_graph_key[1] = -1;  // the graphs hold the capacity in their kernel arguments
  return 0;
}

int ppsd_debug_state(int device, const uint64_t *P, uint32_t n, uint32_t e, uint32_t nwin, uint32_t s,
                     uint64_t *out) {
  if ((!P || !out) && n) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  if (e < 64 || s < 1 || s > 32) return set_err(PPSD_ERR_INVALID_PARAM, "need e >= 64 and 1 <= s <= 32");
  for (uint32_t i = 0; i < n; ++i)
    if (!(P[i] & 1) || P[i] <= ((uint64_t)1 << 32) || P[i] >= ((uint64_t)1 << 62))
      return set_err(PPSD_ERR_INVALID_PARAM, "P[%u] not odd in (2^32, 2^62)", i);
  if (device < 0 || device >= ppsd_device_count()) return set_err(PPSD_ERR_NOT_IMPLEMENTED, "GPU %d not available", device);
  if (n == 0) return 0;
  HIP_TRY(hipSetDevice(device));
  uint64_t *d = nullptr;
  HIP_TRY(hipMalloc(&d, 2 * (size_t)n * sizeof(uint64_t)));
  hipError_t err = hipMemcpy(d, P, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = launch_state(d, n, e, nwin, s, d + n, nullptr);
  if (err == hipSuccess) err = hipMemcpy(out, d + n, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) return set_err(PPSD_ERR_FAILURE, "HIP error: %s", hipGetErrorString(err));
  return 0;
}

