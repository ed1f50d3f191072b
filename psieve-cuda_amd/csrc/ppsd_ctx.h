// ppsd_ctx.h -- the context behind the C ABI, shared by the host-side files
// of libppsd.so (ppsd_api.hip, ppsd_serve.hip).  Not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ppsd.h"
#include "ppsd_batch.h"
#include "ppsd_filter.h"
#include "ppsd_kernels.h"
#include "ppsd_walk.h"

namespace ppsd {

// Records the message for ppsd_last_error (per host thread); returns code.
int set_err(int code, const char *fmt, ...);
uint64_t isqrt64(uint64_t x);
std::vector<uint32_t> host_primes(uint32_t lim);
std::vector<SmallPrime> small_prime_table();
std::vector<uint16_t> pow2_table(const std::vector<SmallPrime> &sm);

enum class Pending { none, primes, range };

// Prime pools of the batched walk: round i is walked from pool i mod kPools
// while rounds i+1 and i+2 are sieved into the others, so that a round's
// sieve has a whole walk to finish in before its own walk needs it.
constexpr int kPools = 3;

// An allocator whose construct() without arguments default-initialises, so
// that resize() before a bulk copy does not zero the new elements first.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <class U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};

// A p-subrange of one unit of a batched search.
struct Piece {
  uint32_t unit;
  uint64_t lo, hi;  // [lo, hi)
  double est;       // expected primes
};

}  // namespace ppsd

#define HIP_TRY(expr)                                                                               \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return ppsd::set_err(PPSD_ERR_FAILURE, "HIP error: %s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

struct ppsd_ctx {
  using Hit = ppsd::Hit;
  using PieceDesc = ppsd::PieceDesc;
  using Task = ppsd::Task;
  using BHit = ppsd::BHit;

  int device = 0;
  hipStream_t stream = nullptr;
  ppsd_params prm{};
  uint32_t batch = 0;
  // the shapes the autotuner found fastest on MI355X (256 CUs); callers that
  // cannot read its cache (the C hosts) get them by default
  int tune_r = 1, tune_u = 16, tune_wg = 64;
  int pre_mode = 0;  // ppsd_set_pretest: 0 automatic, 1 exact, 2 range, 3 odd-low

  // device buffers
  uint64_t *d_P = nullptr;
  uint32_t *d_count = nullptr;             // primes in d_P (range mode)
  unsigned long long *d_counters = nullptr;  // [0] raw hits, [1] kept, [2] prime sum
  Hit *d_hits = nullptr;
  Hit *d_kept = nullptr;
  uint64_t hit_cap = 0;        // capacity in use
  uint64_t hit_alloc = 0;      // allocated entries
  ppsd::SmallPrime *d_small = nullptr;  // odd primes 11 <= q < 32768, padded to a multiple of 64
  uint16_t *d_pow2 = nullptr;           // their powers of two, for the filter
  uint32_t nsmall = 0;
  uint32_t *d_sp = nullptr;     // sieving primes <= qmax
  uint32_t nsp = 0, nsp_small = 0;
  uint64_t qmax = 0;
  uint64_t *d_spinv = nullptr;  // floor((2^64-1)/q) of the LDS-sieved primes d_sp[0..nsp_small)
  std::vector<uint32_t> sp_sample;  // d_sp[0], d_sp[4096], ... (batched walk)
  uint32_t *d_bitmap = nullptr;
  uint32_t *d_blockcount = nullptr;
  unsigned long long *d_blocksum = nullptr;  // per compaction block: prime sum
  uint64_t nbits_max = 0;
  uint64_t *h_stage = nullptr;  // pinned upload staging
  uint8_t *d_cand = nullptr;    // candidate list of a sieve file (ppsd_set_candidates)
  uint64_t cand_row_bytes = 0, cand_b0 = 0;
  uint32_t cand_n0 = 0, cand_rows = 0;

  hipEvent_t ev[6] = {};  // walk begin/end, sieve begin/end, filter begin/end

  // outstanding submission
  ppsd::Pending pending = ppsd::Pending::none;
  uint32_t n_primes_host = 0;  // prime count when known on the host
  bool sieved = false;
  uint64_t range_lo = 0, range_nbits = 0;  // candidates of a range submission
  std::vector<Hit> result;
  bool result_raw = false;
  uint64_t last_primes = 0, last_sum = 0;
  bool have_result = false;

  ppsd_stats st{};

  // batched walk (allocated on first use)
  bool b_ready = false;
  uint32_t cus = 0;  // compute units of the device (read on first use)
  uint64_t b_pool_cap = 0;
  // kPools prime pools: the next rounds are sieved (on stream2) while the
  // current one is walked (on stream)
  uint64_t *b_pool[ppsd::kPools] = {};
  uint32_t *b_pool_count[ppsd::kPools] = {};
  hipStream_t stream2 = nullptr;
  hipEvent_t ev2[ppsd::kPools][2] = {};  // per pool: sieve begin/end on stream2
  uint32_t *h_off[ppsd::kPools] = {};    // per pool, pinned: copies of b_off/b_sums
  unsigned long long *h_sums[ppsd::kPools] = {};
  uint32_t b_piece_cap = 0, b_task_cap = 0;
  PieceDesc *b_pieces = nullptr, *h_pieces = nullptr;
  Task *b_tasks = nullptr;                        // one launch's queue (launch_expand)
  uint32_t *b_ntasks = nullptr, *b_next = nullptr;  // its length, its head
  std::vector<ppsd::TaskClass> b_classes;          // a round's task classes (capacity kept)
  ppsd::TaskClass *b_dclasses = nullptr, *h_classes = nullptr;  // on the device; pinned upload staging
  uint32_t b_class_cap = 0;
  BHit *b_hits = nullptr, *b_kept = nullptr;
  uint64_t b_hit_cap = 0, b_hit_alloc = 0;
  unsigned long long *b_counts = nullptr;  // [0] raw, [1] kept
  uint32_t *b_off[ppsd::kPools] = {};              // per pool: prime offset of each piece (and the count after the last)
  unsigned long long *b_sums[ppsd::kPools] = {};   // per pool: prime sum of each piece
  // a round's sieve (sieve_round): one bitmap of up to kRoundSegs segments
  // per launch, and per pool the pieces and the segment -> piece table
  uint32_t *b_rbitmap = nullptr, *b_rblockcount = nullptr;
  unsigned long long *b_rblocksum = nullptr;
  ppsd::RoundPiece *d_rpieces[ppsd::kPools] = {}, *h_rpieces[ppsd::kPools] = {};
  uint32_t *d_rsegs[ppsd::kPools] = {}, *h_rsegs[ppsd::kPools] = {};
  uint32_t *d_rorder[ppsd::kPools] = {}, *h_rorder[ppsd::kPools] = {};  // per launch, its pieces by descending nq
  uint32_t b_rseg_cap = 0;
  hipGraph_t b_graph[ppsd::kPools] = {};           // per pool: the round graph
  hipGraphExec_t b_exec[ppsd::kPools] = {};
  int b_graph_key[ppsd::kPools] = {-1, -1, -1};
  int b_r = 1, b_u = 16, b_waves = 6;  // defaults: the tuner's choice on MI355X
  uint32_t b_seg_min = 2048, b_seg_max = 65536;
  BHit *b_out = nullptr;  // the batch's results so far, on the device, unit-indexed
  uint64_t b_out_cap = 0, b_out_n = 0;
  std::vector<BHit, ppsd::NoInitAlloc<BHit>> b_result;  // results copied back (batch_flush)
  BHit *h_bstage = nullptr;  // pinned: b_out read back in pieces by ppsd_batch_results
  uint64_t h_bstage_cap = 0;
  // pinned mirror of b_out[0, h_out_cap): each task group's output is copied
  // on copy_stream as soon as it is emitted, beside the next walk, so that
  // the results are on the host when the batch ends (valid once copy_stream
  // has finished)
  BHit *h_out = nullptr;
  uint64_t h_out_cap = 0;
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_emit = nullptr;
  std::vector<uint64_t> b_unit_primes, b_unit_sums;
  uint32_t b_nunits = 0;
};

namespace ppsd {

// Segments of one launch of a round's sieve: a 2^31-bit (256 MiB) bitmap.
constexpr uint32_t kRoundSegs = 1u << 13;

// Copies the device output of the batch so far (b_out) to b_result, after
// the context's stream has finished.
int batch_flush(ppsd_ctx *c);

// What the walk needs of a unit's parameters.
bool walk_domain_ok(const ppsd_params *p);
// Windows of s over [nmin, nmax).
uint32_t windows(const ppsd_params &p);
// Expected raw hits of one prime near lo: every window holds about
// s*(kmax-kmin)/2 candidate values out of p (both signs: twice that).
double hits_per_prime(const ppsd_params &p, uint64_t lo);
// How a unit is walked at kernel shape (r, u): the pre-test form (PRE_EXACT,
// PRE_RANGE or PRE_ODDLO) and the chunk length.  mode 1, 2 and 3 force a
// form (3 falls back to PRE_EXACT when kmax >= 2^32); mode 0 prices them
// with the costs measured on MI355X that choose_form lists.  plain: the
// plain walk kernel, which may shorten the chunk below u and keeps a range
// chunk's states for its re-check; otherwise the persistent kernel, whose
// chunk stays u (it is shared by many units) and which re-walks.
struct WalkForm {
  int pre;
  int u;
};
WalkForm choose_form(const ppsd_params &p, int r, int u, int mode, bool plain);
// Milliseconds between two recorded events (0 on error).
float elapsed(hipEvent_t a, hipEvent_t b);
// Batched walk: buffers (allocated on first use), sieving primes for every
// p < pmax, and the two halves of a round over pieces [b, e) of a piece
// list: enqueue the sieve of the pieces into pool `buf` on stream st
// (ev0/ev1 recorded around it, ev1 after its offsets reached the host);
// then wait for ev1 and walk on c->stream.  walk_round appends the round's
// hits to the device output b_out tagged with their piece's unit (see
// batch_flush), adds to c->b_unit_primes / b_unit_sums, and returns 1 (no
// error) when the pieces overflowed the prime pool (nothing recorded then).
int batch_alloc(ppsd_ctx *c);
int ensure_sieving(ppsd_ctx *c, uint64_t pmax);
int sieve_round(ppsd_ctx *c, const std::vector<Piece> &pieces, size_t b, size_t e, int buf, hipStream_t st,
                hipEvent_t ev0, hipEvent_t ev1);
int walk_round(ppsd_ctx *c, const std::vector<Piece> &pieces, size_t b, size_t e,
               const std::vector<const ppsd_params *> &par, bool raw, int buf, hipEvent_t ev0, hipEvent_t ev1);

}  // namespace ppsd
