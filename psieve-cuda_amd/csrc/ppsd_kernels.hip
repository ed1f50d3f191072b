// ppsd_kernels.hip -- gfx950 kernels of the PPSieve hot path.
//
//   walk_kernel      the n-walk (replaces d_start_ns + d_check_more_ns*,
//                    pps/appcu.cu:1117-1190): one lane per prime, exact hits
//                    compacted in-wave into a device ring
//   filter_kernel    test_factor + try_all_factors on every raw hit
//                    (pps/app.c:150-186, pps/factor_proth.c:220-253)
//   sieve_* / compact_*  segmented Eratosthenes on the device
//                    (replaces sieve.c + main.c:750-777)
//   state_kernel     2^-e mod p then nwin windows; arithmetic check only
#include "ppsd_arith.h"
#include "ppsd_filter.h"
#include "ppsd_kernels.h"
#include "ppsd_walk.h"

namespace ppsd {

// ---------------------------------------------------------------------------
// plain walk: one lane per R primes of one batch
// ---------------------------------------------------------------------------
constexpr uint32_t kPlainStage = 32;  // hits staged per wave in LDS

template <int SIGN, int R, int U, bool S32, int PRE>
__global__ __launch_bounds__(256) void walk_kernel(WalkArgs a) {
  __shared__ Hit stage[4][kPlainStage];
  __shared__ uint32_t fill[4];
  const uint32_t cnt = a.count_d ? *a.count_d : a.count;
  const uint32_t count = cnt < a.cap ? cnt : a.cap;
  const uint32_t first = (blockIdx.x * blockDim.x + threadIdx.x) * R;
  if (first >= count) return;
  LanePrimes<R> L;
  LaneStates<R> S;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    L.valid[r] = first + r < count;
    L.p[r] = a.P[L.valid[r] ? first + r : count - 1];
  }
  const WindowRule rule{a.kmin, a.kmax, a.nmax, a.s, a.kmax << (a.s - 1)};
  start_states<SIGN, R>(L, S, a.nmin, a.s, rule.bound);
  const uint32_t wave = threadIdx.x >> 6;
  fill[wave] = 0;
  const StagedSink<RingSink, kPlainStage> sink{RingSink{a.hits, a.hit_count, a.hit_cap}, stage[wave], &fill[wave]};
  walk_windows<SIGN, R, U, S32, PRE, PRE == PRE_RANGE>(L, S, a.nmin, a.nwin, rule, sink);
  sink.flush();
}

template <int SIGN, int R, bool S32, int PRE>
static hipError_t launch_walk_u(const WalkArgs &a, int u, uint32_t grid, uint32_t block, hipStream_t st) {
  switch (u) {
    case 4: hipLaunchKernelGGL((walk_kernel<SIGN, R, 4, S32, PRE>), dim3(grid), dim3(block), 0, st, a); break;
    case 8: hipLaunchKernelGGL((walk_kernel<SIGN, R, 8, S32, PRE>), dim3(grid), dim3(block), 0, st, a); break;
    case 16: hipLaunchKernelGGL((walk_kernel<SIGN, R, 16, S32, PRE>), dim3(grid), dim3(block), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int SIGN, bool S32, int PRE>
static hipError_t launch_walk_r(const WalkArgs &a, int r, int u, uint32_t grid, uint32_t block, hipStream_t st) {
  switch (r) {
    case 1: return launch_walk_u<SIGN, 1, S32, PRE>(a, u, grid, block, st);
    case 2: return launch_walk_u<SIGN, 2, S32, PRE>(a, u, grid, block, st);
    default: return hipErrorInvalidValue;
  }
}

template <bool S32, int PRE>
static hipError_t launch_walk_s(const WalkArgs &a, int sign, int r, int u, uint32_t grid, uint32_t block,
                                hipStream_t st) {
  switch (sign) {
    case 1: return launch_walk_r<1, S32, PRE>(a, r, u, grid, block, st);
    case -1: return launch_walk_r<-1, S32, PRE>(a, r, u, grid, block, st);
    case 0: return launch_walk_r<0, S32, PRE>(a, r, u, grid, block, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_walk(const WalkArgs &a, int sign, int r, int u, uint32_t grid, uint32_t block, hipStream_t st) {
  if (a.s < 1 || a.s > 32) return hipErrorInvalidValue;
  if (a.pre == PRE_RANGE)
    return a.s == 32 ? launch_walk_s<true, PRE_RANGE>(a, sign, r, u, grid, block, st)
                     : launch_walk_s<false, PRE_RANGE>(a, sign, r, u, grid, block, st);
  if (a.pre == PRE_ODDLO) {
    if (a.kmax >> 32) return hipErrorInvalidValue;
    return a.s == 32 ? launch_walk_s<true, PRE_ODDLO>(a, sign, r, u, grid, block, st)
                     : launch_walk_s<false, PRE_ODDLO>(a, sign, r, u, grid, block, st);
  }
  return a.s == 32 ? launch_walk_s<true, PRE_EXACT>(a, sign, r, u, grid, block, st)
                   : launch_walk_s<false, PRE_EXACT>(a, sign, r, u, grid, block, st);
}

// ---------------------------------------------------------------------------
// filter: test_factor (pps/app.c:159-186) on every raw hit
// ---------------------------------------------------------------------------
struct PlainRule {
  uint32_t kstep, koffset, nmin, nmax;
  __device__ void operator()(const Hit &, uint32_t &ks, uint32_t &ko, uint32_t &n0, uint32_t &n1) const {
    ks = kstep;
    ko = koffset;
    n0 = nmin;
    n1 = nmax;
  }
};

__global__ __launch_bounds__(kFilterThreads) void filter_kernel(FilterArgs a) {
  __shared__ __align__(16) uint16_t pass_pow[kPowPassBytes / sizeof(uint16_t)];
  const unsigned long long raw = *a.hit_count;
  const uint64_t nraw = raw < a.hit_cap ? raw : a.hit_cap;
  filter_hits(a.hits, nraw, a.kept, a.kept_count, a.kept_cap, a.small_primes, a.pow2, a.nsmall,
              PlainRule{a.kstep, a.koffset, a.nmin, a.nmax}, pass_pow);
}

// With a sieve file's candidate list the rule is membership alone
// (pps/app.c:181-184): one hit per thread.
__global__ __launch_bounds__(256) void filter_cand_kernel(FilterArgs a) {
  const unsigned long long raw = *a.hit_count;
  const uint64_t nraw = raw < a.hit_cap ? raw : a.hit_cap;
  // wave-uniform loop (the stride is a whole number of waves), so that the
  // kept lanes of a wave reserve their slots together
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u);
  for (uint64_t h0 = first; h0 < nraw; h0 += stride) {
    const uint64_t h = h0 + __lane_id();
    Hit x{};
    bool keep = h < nraw;
    if (keep) {
      x = a.hits[h];
      const uint64_t b = x.k / a.kstep;
      keep = x.k == (uint64_t)a.kstep * b + a.koffset && x.n >= a.nmin && x.n < a.nmax;
      if (keep) {
        const uint64_t row = (uint64_t)x.n - a.cand_n0;
        const uint64_t bit = b - a.cand_b0;
        keep = x.n >= a.cand_n0 && row < a.cand_rows && b >= a.cand_b0 && bit < a.cand_row_bytes * 8 &&
               ((a.cand[row * a.cand_row_bytes + bit / 8] >> (bit % 8)) & 1u);
      }
    }
    const uint64_t keepmask = __ballot(keep);
    const unsigned long long idx = wave_slot(keepmask, a.kept_count);
    if (keep && idx < a.kept_cap) a.kept[idx] = x;
  }
}

hipError_t launch_filter(const FilterArgs &a, uint32_t grid, hipStream_t st) {
  if (a.cand)
    hipLaunchKernelGGL(filter_cand_kernel, dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(filter_kernel, dim3(grid), dim3(kFilterThreads), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// segmented sieve of Eratosthenes over odd candidates
// ---------------------------------------------------------------------------

// First odd multiple of q that is >= max(q^2, from); from is odd.
__device__ __forceinline__ uint64_t first_multiple(uint64_t from, uint32_t q) {
  const uint64_t qq = (uint64_t)q * q;
  if (qq >= from) return qq;
  const uint32_t r = (uint32_t)(from % q);
  uint64_t m = from + (r ? q - r : 0);
  if ((m & 1) == 0) m += q;
  return m;
}

// The same, with the remainder from inv = floor((2^64-1)/q): the estimate
// hi64(from*inv) is the quotient or one less, so one subtraction finishes.
__device__ __forceinline__ uint64_t first_multiple_inv(uint64_t from, uint32_t q, uint64_t inv) {
  const uint64_t qq = (uint64_t)q * q;
  if (qq >= from) return qq;
  uint64_t r = from - __umul64hi(from, inv) * q;
  r -= r >= q ? q : 0;
  uint64_t m = from + (r ? q - r : 0);
  if ((m & 1) == 0) m += q;
  return m;
}

// One workgroup per segment of kSegBits candidates.  Primes below 1024 are
// marked cooperatively by the whole workgroup (each has many multiples in a
// segment); the rest of q[0..nq_small) one prime per thread.  Every bitmap
// word of the segment is written, tail bits past nbits set.  BS threads:
// 1024 where the sieve has the chip to itself; 256 (one wave per SIMD) where
// it runs beside a persistent walk that leaves one wave slot per SIMD free,
// which a 1024-thread workgroup (four waves per SIMD) could never get.
// The body of one segment: candidates seg_lo + 2b, b < nb (nb <= kSegBits),
// marked in LDS (seg) with q[0..nq_small) (qinv: their reciprocals, or
// null); then nw_out words go to out, every bit past nb set.
__device__ __forceinline__ void sieve_segment(uint32_t *seg, uint64_t seg_lo, uint32_t nb, const uint32_t *qs,
                                              const uint64_t *qinv, uint32_t nq_small, uint32_t *out,
                                              uint32_t nw_out) {
  for (uint32_t w = threadIdx.x; w < kSegWords; w += blockDim.x) seg[w] = 0;
  __syncthreads();
  uint32_t i = 0;
  for (; i < nq_small && qs[i] < 1024u; ++i) {
    const uint32_t q = qs[i];
    const uint64_t m = qinv ? first_multiple_inv(seg_lo, q, qinv[i]) : first_multiple(seg_lo, q);
    const uint64_t off = (m - seg_lo) >> 1;
    if (off >= nb) continue;
    for (uint64_t b = off + (uint64_t)threadIdx.x * q; b < nb; b += (uint64_t)blockDim.x * q)
      atomicOr(&seg[b >> 5], 1u << (b & 31));
  }
  for (uint32_t t = i + threadIdx.x; t < nq_small; t += blockDim.x) {
    const uint32_t q = qs[t];
    const uint64_t m = qinv ? first_multiple_inv(seg_lo, q, qinv[t]) : first_multiple(seg_lo, q);
    for (uint64_t b = (m - seg_lo) >> 1; b < nb; b += q) atomicOr(&seg[b >> 5], 1u << (b & 31));
  }
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < nw_out; w += blockDim.x) {
    uint32_t v = seg[w];
    const uint32_t bb = w * 32;
    if (bb >= nb) v = ~0u;
    else if (bb + 32 > nb) v |= ~((1u << (nb - bb)) - 1u);
    out[w] = v;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void sieve_small_kernel(SieveArgs a) {
  __shared__ uint32_t seg[kSegWords];
  const uint64_t b0 = (uint64_t)blockIdx.x * kSegBits;
  const uint64_t rem = a.nbits - b0;
  const uint32_t nb = rem < kSegBits ? (uint32_t)rem : kSegBits;
  sieve_segment(seg, a.lo + 2 * b0, nb, a.q, a.qinv, a.nq_small, a.bitmap + b0 / 32, (nb + 31) / 32);
}

// The kernels of a round sieve run on a second stream beside the batched
// walk, whose waves keep every SIMD's VALU issue busy.  A SIMD issues from
// the highest-priority ready wave, the oldest among equals: at the default
// priority a sieve wave next to the walk's older waves issues almost never,
// and the sieve finishes only after the walk has, in the gap before the next
// walk.  At the highest priority it issues whenever it is ready; it spends
// most of its time waiting on LDS and memory atomics, so the instructions it
// takes from the walk are few.
__device__ __forceinline__ void beside_walk_priority() { __builtin_amdgcn_s_setprio(3); }

// One workgroup per segment of a round launch: the segment's piece from
// seg_piece; all kSegWords words are written (a piece's last segment is
// padded with composites), so the compaction can run over whole blocks.
template <int BS>
__global__ __launch_bounds__(BS) void round_sieve_small_kernel(RoundSieveArgs a) {
  __shared__ uint32_t seg[kSegWords];
  beside_walk_priority();
  for (uint32_t g = blockIdx.x; g < a.nseg; g += gridDim.x) {  // workgroup-uniform
    const RoundPiece P = a.pieces[a.seg_piece[g]];
    const uint64_t b0 = (uint64_t)g * kSegBits - P.bit0;  // within the piece
    const uint64_t rem = P.nbits - b0;
    const uint32_t nb = rem < kSegBits ? (uint32_t)rem : kSegBits;
    sieve_segment(seg, P.lo + 2 * b0, nb, a.q, a.qinv, P.nq_small, a.bitmap + (uint64_t)g * kSegWords, kSegWords);
    __syncthreads();  // seg is read out before the next segment clears it
  }
}

// The large primes of a round launch, one per thread: q = q[t] for t in
// [qlo, qhi), marked in every piece of the launch that sieves with it (those
// whose nq exceeds t: a prefix of by_nq).  The reciprocal floor((2^64-1)/q)
// is formed once per prime, so that each piece's first multiple costs a
// multiply-high and one correction instead of a 64-bit remainder.
__global__ __launch_bounds__(256) void round_sieve_large_kernel(RoundSieveArgs a) {
  beside_walk_priority();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)a.qlo + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.qhi; t += stride) {
    const uint32_t q = a.q[t];
    const uint64_t inv = ~0ull / q;
    for (uint32_t j = a.p0; j < a.p1; ++j) {
      const RoundPiece P = a.pieces[a.by_nq[j]];
      if (P.nq <= t) break;
      if (t < P.nq_small) continue;  // marked in LDS for this piece
      const uint64_t m = first_multiple_inv(P.lo, q, inv);
      for (uint64_t b = (m - P.lo) >> 1; b < P.nbits; b += q) {
        const uint64_t bit = P.bit0 + b;
        atomicOr(&a.bitmap[bit >> 5], 1u << (bit & 31));
      }
    }
  }
}

// Primes q[nq_small..nq): at most a few multiples per chunk each; one thread
// per prime, marks with global atomics (after sieve_small_kernel wrote every
// word of the chunk).
__global__ __launch_bounds__(256) void sieve_large_kernel(SieveArgs a) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t t = a.nq_small + blockIdx.x * blockDim.x + threadIdx.x; t < a.nq; t += stride) {
    const uint32_t q = a.q[t];
    const uint64_t m = first_multiple(a.lo, q);
    for (uint64_t b = (m - a.lo) >> 1; b < a.nbits; b += q) atomicOr(&a.bitmap[b >> 5], 1u << (b & 31));
  }
}

hipError_t launch_sieve(const SieveArgs &a, hipStream_t st) {
  const uint32_t nseg = (uint32_t)((a.nbits + kSegBits - 1) / kSegBits);
  // 1024 threads: a segment's block walks every LDS-sieved prime (up to 2^21)
  if (a.block == 256)
    hipLaunchKernelGGL(sieve_small_kernel<256>, dim3(nseg), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(sieve_small_kernel<1024>, dim3(nseg), dim3(1024), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.nq > a.nq_small) {
    uint32_t grid = (a.nq - a.nq_small + 255) / 256;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(sieve_large_kernel, dim3(grid), dim3(256), 0, st, a);
    e = hipGetLastError();
  }
  return e;
}

// ---------------------------------------------------------------------------
// compaction of a chunk bitmap into an ascending prime list
// ---------------------------------------------------------------------------

// inclusive prefix sum over a 256-thread workgroup; returns the exclusive one
// and the total through `total`
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t wsum[4];
  const uint32_t lane = __lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t j = 0; j < wid; ++j) before += wsum[j];
  total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();
  return before + x - v;
}

// count_kernel and scan_kernel also end every round sieve, so they take the
// round sieve's wave priority (the plain path's sieve uses them too).
__global__ __launch_bounds__(256) void count_kernel(const uint32_t *bitmap, uint32_t nwords, uint32_t *blockcount) {
  beside_walk_priority();
  const uint32_t idx = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t c = idx < nwords ? (uint32_t)__popc(~bitmap[idx]) : 0u;
  uint32_t total;
  block_exclusive_scan(c, total);
  if (threadIdx.x == 0) blockcount[blockIdx.x] = total;
}

// Exclusive scan of blockcount in place, offset by the running count in
// *count_d, which is then advanced by the chunk's total.  One workgroup.
__global__ __launch_bounds__(256) void scan_kernel(uint32_t *blockcount, uint32_t nblocks, uint32_t *count_d) {
  beside_walk_priority();
  const uint32_t per = (nblocks + 255) / 256;
  const uint32_t beg = threadIdx.x * per;
  const uint32_t end = beg + per < nblocks ? beg + per : nblocks;
  uint32_t local = 0;
  for (uint32_t j = beg; j < end; ++j) local += blockcount[j];
  uint32_t total;
  const uint32_t ex = block_exclusive_scan(local, total);
  const uint32_t base = *count_d;
  uint32_t run = base + ex;
  for (uint32_t j = beg; j < end; ++j) {
    const uint32_t c = blockcount[j];
    blockcount[j] = run;
    run += c;
  }
  __syncthreads();
  if (threadIdx.x == 0) *count_d = base + total;
}

template <typename T>
__global__ __launch_bounds__(256) void compact_kernel(const uint32_t *bitmap, uint32_t nwords, uint64_t lo,
                                                      const uint32_t *blockoff, T *out, uint32_t cap,
                                                      unsigned long long *blocksum) {
  const uint32_t idx = blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t w = idx < nwords ? ~bitmap[idx] : 0u;
  uint32_t total;
  uint32_t pos = blockoff[blockIdx.x] + block_exclusive_scan((uint32_t)__popc(w), total);
  unsigned long long s = 0;
  while (w) {
    const uint32_t bit = (uint32_t)__builtin_ctz(w);
    w &= w - 1;
    const uint64_t x = lo + 2 * ((uint64_t)idx * 32 + bit);
    if (pos < cap) out[pos] = (T)x;  // the count still advances: the host sees an overflow
    ++pos;
    s += x;
  }
  if (blocksum) {
    // the block's prime sum, stored for sum_blocks_kernel (atomics on one
    // address from every wave serialised this kernel)
    __shared__ unsigned long long wsum[4];
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) blocksum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  }
}

// Adds the blocks' prime sums to *sum_d (one workgroup).
__global__ __launch_bounds__(256) void sum_blocks_kernel(const unsigned long long *blocksum, uint32_t n,
                                                         unsigned long long *sum_d) {
  __shared__ unsigned long long part[256];
  unsigned long long s = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += blocksum[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t h = 128; h >= 1; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *sum_d += part[0];
}

template <typename T>
static hipError_t launch_compact(const uint32_t *bitmap, uint64_t lo, uint64_t nbits, uint32_t *blockcount, T *out,
                                 uint32_t cap, uint32_t *count_d, unsigned long long *blocksum,
                                 unsigned long long *sum_d, hipStream_t st) {
  const uint32_t nwords = (uint32_t)((nbits + 31) / 32);
  const uint32_t nblocks = (nwords + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(count_kernel, dim3(nblocks), dim3(256), 0, st, bitmap, nwords, blockcount);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, st, blockcount, nblocks, count_d);
  hipLaunchKernelGGL((compact_kernel<T>), dim3(nblocks), dim3(256), 0, st, bitmap, nwords, lo,
                     (const uint32_t *)blockcount, out, cap, sum_d ? blocksum : nullptr);
  if (sum_d) hipLaunchKernelGGL(sum_blocks_kernel, dim3(1), dim3(256), 0, st, blocksum, nblocks, sum_d);
  return hipGetLastError();
}

hipError_t launch_compact64(const uint32_t *bitmap, uint64_t lo, uint64_t nbits, uint32_t *blockcount,
                            unsigned long long *blocksum, uint64_t *out, uint32_t cap, uint32_t *count_d,
                            unsigned long long *sum_d, hipStream_t st) {
  return launch_compact<uint64_t>(bitmap, lo, nbits, blockcount, out, cap, count_d, blocksum, sum_d, st);
}

hipError_t launch_compact32(const uint32_t *bitmap, uint64_t lo, uint64_t nbits, uint32_t *blockcount, uint32_t *out,
                            uint32_t cap, uint32_t *count_d, hipStream_t st) {
  return launch_compact<uint32_t>(bitmap, lo, nbits, blockcount, out, cap, count_d, nullptr, nullptr, st);
}

// compact_kernel over a round launch's bitmap: a block's candidates belong
// to the piece of its segment.
__global__ __launch_bounds__(256) void round_compact_kernel(RoundSieveArgs a, const uint32_t *blockoff, uint64_t *out,
                                                            uint32_t cap, unsigned long long *blocksum) {
  beside_walk_priority();
  const uint32_t blk = blockIdx.x;
  const RoundPiece P = a.pieces[a.seg_piece[blk / kBlocksPerSeg]];
  const uint64_t idx = (uint64_t)blk * kScanBlock + threadIdx.x;
  uint32_t w = ~a.bitmap[idx];
  uint32_t total;
  uint32_t pos = blockoff[blk] + block_exclusive_scan((uint32_t)__popc(w), total);
  unsigned long long s = 0;
  while (w) {
    const uint32_t bit = (uint32_t)__builtin_ctz(w);
    w &= w - 1;
    const uint64_t x = P.lo + 2 * (idx * 32 + bit - P.bit0);
    if (pos < cap) out[pos] = x;  // the count still advances: the host sees an overflow
    ++pos;
    s += x;
  }
  __shared__ unsigned long long wsum[4];
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if (__lane_id() == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) blocksum[blk] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Per piece of a round launch (one wave each): the pool offset of its first
// prime and the sum of its primes, from the blocks of its segments.
__global__ __launch_bounds__(64) void round_piece_totals_kernel(RoundSieveArgs a, const uint32_t *blockoff,
                                                                const unsigned long long *blocksum, uint32_t nblocks,
                                                                const uint32_t *count_d, uint32_t *off,
                                                                unsigned long long *psum) {
  beside_walk_priority();
  const uint32_t i = a.p0 + blockIdx.x;
  const RoundPiece P = a.pieces[i];
  const uint32_t first = (uint32_t)(P.bit0 / (kScanBlock * 32));
  const uint32_t nblk = (uint32_t)((P.nbits + kSegBits - 1) / kSegBits) * kBlocksPerSeg;
  unsigned long long s = 0;
  for (uint32_t j = __lane_id(); j < nblk; j += 64) s += blocksum[first + j];
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if (__lane_id() == 0) {
    off[i] = first < nblocks ? blockoff[first] : *count_d;  // an empty piece at the end: the final count
    psum[i] = s;
  }
}

hipError_t launch_round_sieve(const RoundSieveArgs &a, uint32_t *blockcount, unsigned long long *blocksum,
                              uint64_t *out, uint32_t cap, uint32_t *count_d, uint32_t *off,
                              unsigned long long *psum, hipStream_t st) {
  const uint32_t nblocks = a.nseg * kBlocksPerSeg;
  if (a.nseg) {
    // a.grid caps both sieve launches (one workgroup per segment or per 256
    // large primes otherwise): beside a persistent walk, a launch of
    // thousands of workgroups would take every slot the walk's end frees and
    // hold off the next walk until it drained
    const uint32_t gs = a.grid ? std::min(a.nseg, a.grid) : a.nseg;
    if (a.block == 256)
      hipLaunchKernelGGL(round_sieve_small_kernel<256>, dim3(gs), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL(round_sieve_small_kernel<1024>, dim3(gs), dim3(1024), 0, st, a);
    if (a.qhi > a.qlo) {
      const uint32_t gl = (a.qhi - a.qlo + 255) / 256;
      hipLaunchKernelGGL(round_sieve_large_kernel, dim3(a.grid ? std::min(gl, a.grid) : gl), dim3(256), 0, st, a);
    }
    hipLaunchKernelGGL(count_kernel, dim3(nblocks), dim3(256), 0, st, a.bitmap, nblocks * kScanBlock, blockcount);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, st, blockcount, nblocks, count_d);
    hipLaunchKernelGGL(round_compact_kernel, dim3(nblocks), dim3(256), 0, st, a, (const uint32_t *)blockcount, out,
                       cap, blocksum);
  }
  if (a.p1 > a.p0)
    hipLaunchKernelGGL(round_piece_totals_kernel, dim3(a.p1 - a.p0), dim3(64), 0, st, a,
                       (const uint32_t *)blockcount, (const unsigned long long *)blocksum, nblocks,
                       (const uint32_t *)count_d, off, psum);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void zero_u32_kernel(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void zero_u64_kernel(unsigned long long *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// 2^-e mod p followed by nwin windows of s: the arithmetic alone, for checks
// against closed forms at any p.  The state is returned as the walk holds it
// (no final reduction), so the checks also see that it stays canonical.
__global__ __launch_bounds__(256) void state_kernel(const uint64_t *P, uint32_t n, uint32_t e, uint32_t nwin,
                                                    uint32_t s, uint64_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = P[i];
  const uint64_t pinv = inv64(p);
  const uint32_t pinv_sh = step_multiplier(pinv, s);
  uint64_t K = inv2pow(p, pinv, e);
  const uint32_t zero = opaque_zero();
  if (s == 32 && fused_step_ok(p, s))
    for (uint32_t w = 0; w < nwin; ++w) K = step_fused<true>(K, p, pinv_sh, s, zero);
  else if (fused_step_ok(p, s))
    for (uint32_t w = 0; w < nwin; ++w) K = step_fused<false>(K, p, pinv_sh, s, zero);
  else if (s == 32)
    for (uint32_t w = 0; w < nwin; ++w) K = step<true>(K, p, pinv_sh, s, zero);
  else
    for (uint32_t w = 0; w < nwin; ++w) K = step<false>(K, p, pinv_sh, s, zero);
  out[i] = K;
}

hipError_t launch_state(const uint64_t *P, uint32_t n, uint32_t e, uint32_t nwin, uint32_t s, uint64_t *out,
                        hipStream_t st) {
  hipLaunchKernelGGL(state_kernel, dim3((n + 255) / 256), dim3(256), 0, st, P, n, e, nwin, s, out);
  return hipGetLastError();
}

}  // namespace ppsd
