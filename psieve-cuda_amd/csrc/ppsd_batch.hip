// ppsd_batch.hip -- many independent work units in one persistent launch.
//
// A work unit is a full search (its own p-range, k-range, n-range, sign and
// window).  Units are cut into pieces (p-subranges) that fit the device prime
// pool; a round sieves its pieces into the pool, then a device work queue of
// tasks -- (piece, block of 64*R primes, segment of windows) -- is drained by
// a persistent kernel whose waves pull tasks with one atomic each until the
// queue is empty.  Splitting n-ranges into segments is what keeps the whole
// chip busy on units with few primes; the segment start state is recomputed
// with one 2^-e powmod per lane.  The round's zeroing, walk and filter are
// captured once as a hipGraph and replayed for every round.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ppsd.h"
#include "ppsd_batch.h"
#include "ppsd_filter.h"
#include "ppsd_walk.h"

namespace ppsd {

// Raw-hit ring of the batched walk: 32-byte records tagged with the piece.
struct PieceSink {
  using Rec = BHit;
  BHit *hits;
  unsigned long long *count;
  uint64_t cap;
  uint32_t piece;
  __device__ __forceinline__ BHit make(uint64_t p, uint64_t k, uint32_t n, int32_t c) const {
    BHit h;
    h.p = p;
    h.k = k;
    h.n = n;
    h.c = c;
    h.piece = piece;
    h.pad = 0;
    return h;
  }
};

// Hits staged per wave in LDS across tasks (each record carries its piece);
// flushed when the stage fills and when the wave leaves.
constexpr uint32_t kBatchStage = 64;

template <int SIGN, int R, int U, bool S32, int PRE>
__device__ __forceinline__ void run_task(const BatchArgs &a, const PieceDesc &P, const Task &T, BHit *stage,
                                         uint32_t *fill) {
  LanePrimes<R> L;
  LaneStates<R> S;
  const uint32_t first = T.first + __lane_id() * R;
  const uint32_t end = T.first + T.count;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    L.valid[r] = first + r < end;
    L.p[r] = a.pool[P.prime_off + (L.valid[r] ? first + r : T.first)];
  }
  const uint32_t n0 = P.nmin + T.w0 * P.s;
  const WindowRule rule{P.kmin, P.kmax, P.nmax, P.s, P.kmax << (P.s - 1)};
  start_states<SIGN, R>(L, S, n0, P.s, rule.bound);
  const StagedSink<PieceSink, kBatchStage> sink{PieceSink{a.hits, a.hit_count, a.hit_cap, T.piece}, stage, fill};
  // A chunk's states are kept for its re-check at U = 4 (about 3 of the 7
  // instructions of a re-walked window saved; at r = 1 68 VGPRs, still 7
  // waves per SIMD) and re-walked from the snapshot at U = 8 and 16, where U kept
  // states would cost the kernel, which holds every form, resident waves.
  walk_windows<SIGN, R, U, S32, PRE, (U <= 4)>(L, S, n0, T.nw, rule, sink);
}

template <int R, int U, int PRE>
__device__ __forceinline__ void run_form(const BatchArgs &a, const PieceDesc &P, const Task &T, BHit *stage,
                                         uint32_t *fill) {
  const bool s32 = P.s == 32;
  if (P.sign == 1) {
    if (s32) run_task<1, R, U, true, PRE>(a, P, T, stage, fill);
    else run_task<1, R, U, false, PRE>(a, P, T, stage, fill);
  } else if (P.sign == -1) {
    if (s32) run_task<-1, R, U, true, PRE>(a, P, T, stage, fill);
    else run_task<-1, R, U, false, PRE>(a, P, T, stage, fill);
  } else {
    if (s32) run_task<0, R, U, true, PRE>(a, P, T, stage, fill);
    else run_task<0, R, U, false, PRE>(a, P, T, stage, fill);
  }
}

// Every wave pulls tasks until the queue is empty; all waves reach the exit,
// flushing their staged hits first.
template <int R, int U>
__global__ __launch_bounds__(256) void persist_walk_kernel(BatchArgs a) {
  __shared__ BHit stage[4][kBatchStage];
  __shared__ uint32_t fill[4];
  const uint32_t wave = threadIdx.x >> 6;
  fill[wave] = 0;
  const uint32_t ntasks = *a.ntasks;
  for (;;) {
    uint32_t t = 0;
    if (__lane_id() == 0) t = atomicAdd(a.next_task, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
    if (t >= ntasks) {
      const StagedSink<PieceSink, kBatchStage> sink{PieceSink{a.hits, a.hit_count, a.hit_cap, 0}, stage[wave],
                                                    &fill[wave]};
      sink.flush();
      return;
    }
    const Task T = a.tasks[t];
    if (T.nw == 0) continue;  // an unused part of a cut tail task
    const PieceDesc P = a.pieces[T.piece];
    if (P.pre == PRE_RANGE) run_form<R, U, PRE_RANGE>(a, P, T, stage[wave], &fill[wave]);
    else if (P.pre == PRE_ODDLO) run_form<R, U, PRE_ODDLO>(a, P, T, stage[wave], &fill[wave]);
    else run_form<R, U, PRE_EXACT>(a, P, T, stage[wave], &fill[wave]);
  }
}

// test_factor on every raw hit of the round, with the rule of the hit's piece.
struct PieceRule {
  const PieceDesc *pieces;
  __device__ void operator()(const BHit &x, uint32_t &ks, uint32_t &ko, uint32_t &n0, uint32_t &n1) const {
    const PieceDesc &P = pieces[x.piece];
    ks = P.kstep;
    ko = P.koffset;
    n0 = P.nmin;
    n1 = P.nmax;
  }
};

__global__ __launch_bounds__(kFilterThreads) void batch_filter_kernel(BatchArgs a) {
  __shared__ __align__(16) uint16_t pass_pow[kPowPassBytes / sizeof(uint16_t)];
  const unsigned long long raw = *a.hit_count;
  const uint64_t nraw = raw < a.hit_cap ? raw : a.hit_cap;
  filter_hits(a.hits, nraw, a.kept, a.kept_count, a.kept_cap, a.small_primes, a.pow2, a.nsmall,
              PieceRule{a.pieces}, pass_pow);
}

// Appends n hits of a ring to the batch's output on the device, with the
// round-local piece index of each replaced by the caller's unit index.
__global__ __launch_bounds__(256) void emit_kernel(const BHit *src, uint64_t n, const PieceDesc *pieces, BHit *dst) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    BHit h = src[i];
    h.piece = pieces[h.piece].unit;
    dst[i] = h;
  }
}

// Task t of the round's order, from its class (binary search over t_begin).
__device__ __forceinline__ Task task_of(const TaskClass *cl, uint32_t ncl, uint64_t t) {
  uint32_t lo = 0, hi = ncl;  // cl[lo].t_begin <= t < cl[hi].t_begin (hi = ncl: past the end)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cl[mid].t_begin <= t) lo = mid;
    else hi = mid;
  }
  const TaskClass &C = cl[lo];
  const uint64_t local = t - C.t_begin;
  const uint32_t i = (uint32_t)(local / C.nseg), j = (uint32_t)(local - (uint64_t)i * C.nseg);
  return Task{C.piece, C.f0 + i * C.fstep, C.count, C.w0 + j * C.nw, C.nw};
}

__global__ __launch_bounds__(256) void expand_kernel(ExpandArgs a) {
  const uint64_t n = a.t1 - a.t0, head = n - a.tail, total = head + (uint64_t)a.tail * a.parts;
  const uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (q == 0) *a.nout = (uint32_t)total;
  if (q >= total) return;
  if (q < head) {
    a.out[q] = task_of(a.classes, a.nclasses, a.t0 + q);
    return;
  }
  // part `part` of tail task `i`: windows [nw*part/parts_i, nw*(part+1)/parts_i)
  const uint64_t u = q - head;
  const uint32_t i = (uint32_t)(u % a.tail), part = (uint32_t)(u / a.tail);
  Task T = task_of(a.classes, a.nclasses, a.t0 + head + i);
  const uint32_t parts = max(1u, min(a.parts, T.nw / a.tail_min_windows));
  if (part >= parts) {
    T.nw = 0;
  } else {
    const uint32_t w0 = (uint32_t)((uint64_t)T.nw * part / parts), w1 = (uint32_t)((uint64_t)T.nw * (part + 1) / parts);
    T.w0 += w0;
    T.nw = w1 - w0;
  }
  a.out[q] = T;
}

hipError_t launch_expand(const ExpandArgs &a, hipStream_t st) {
  const uint64_t total = (a.t1 - a.t0 - a.tail) + (uint64_t)a.tail * a.parts;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, (total + 255) / 256);
  hipLaunchKernelGGL(expand_kernel, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <int R, int U>
static hipError_t launch_persist_u(const BatchArgs &a, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL((persist_walk_kernel<R, U>), dim3(grid), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_persist(const BatchArgs &a, int r, int u, uint32_t grid, hipStream_t st) {
  if (r == 1) {
    if (u == 4) return launch_persist_u<1, 4>(a, grid, st);
    if (u == 8) return launch_persist_u<1, 8>(a, grid, st);
    if (u == 16) return launch_persist_u<1, 16>(a, grid, st);
  } else if (r == 2) {
    if (u == 4) return launch_persist_u<2, 4>(a, grid, st);
    if (u == 8) return launch_persist_u<2, 8>(a, grid, st);
    if (u == 16) return launch_persist_u<2, 16>(a, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_filter(const BatchArgs &a, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL(batch_filter_kernel, dim3(grid), dim3(kFilterThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_emit(const BHit *src, uint64_t n, const PieceDesc *pieces, BHit *dst, hipStream_t st) {
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
  if (blocks) hipLaunchKernelGGL(emit_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, src, n, pieces, dst);
  return hipGetLastError();
}

}  // namespace ppsd
