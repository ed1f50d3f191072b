// ppsd_walk.h -- the window loop shared by the plain and the persistent walk
// kernels: pre-test, exact test, in-wave hit compaction, chunked stepping.
#pragma once
#include "ppsd_arith.h"
#include "ppsd_kernels.h"

namespace ppsd {

// What a window test needs besides the state.
struct WindowRule {
  uint64_t kmin, kmax;
  uint32_t nmax;   // exclusive
  uint32_t s;      // window, 1..32
  uint64_t bound;  // kmax * 2^(s-1): no hit of a window exceeds it
};

// Pre-test forms.  PRE_EXACT: the odd part of the state (ctz, shift,
// compare) is <= kmax.  PRE_RANGE: the state's high word is <= that of
// kmax * 2^(s-1) (both signs: or >= that of p minus it), as a running
// minimum (maximum) over the chunk -- a necessary condition that also fires
// on states that are merely small, which costs a re-walk of the chunk; the
// host picks it when that is rare, i.e. when kmax * 2^(s-1) is a tiny
// fraction of p.  PRE_ODDLO (kmax < 2^32 only): the
// low word of the odd part, minimised over the chunk, is <= kmax -- see
// odd_low.
enum { PRE_EXACT = 0, PRE_RANGE = 1, PRE_ODDLO = 2 };

// Low word of the odd part of K: v_ffbl_b32, then v_alignbit_b32 takes bits
// i..i+31 of K.  A hit K = k*2^i (k odd, k <= kmax < 2^32, i < s <= 32)
// gives exactly k, so "the chunk's minimum is <= kmax" is necessary for a
// hit in the chunk; it also fires on about kmax/2^32 of other states (whose
// odd part is large but ends in a small low word), and on some states with
// a zero low word, which hold no hit (i >= 32 >= s).  Two instructions per
// state and one v_min3_u32 per two states, against the exact form's three
// instructions (v_ffbl_b32, 64-bit shift, 64-bit compare) per state: in a
// stream of multiplies every VALU instruction takes one issue slot, so the
// instruction count is what matters.
__device__ __forceinline__ uint32_t odd_low(uint64_t K) {
  return __builtin_amdgcn_alignbit(hi32(K), lo32(K), ffbl(lo32(K)));
}

// Raw-hit ring of the plain walk: 24-byte records.
struct RingSink {
  using Rec = Hit;
  Hit *hits;
  unsigned long long *count;
  uint64_t cap;
  __device__ __forceinline__ Hit make(uint64_t p, uint64_t k, uint32_t n, int32_t c) const {
    Hit h;
    h.p = p;
    h.k = k;
    h.n = n;
    h.c = c;
    return h;
  }
};

// Hits staged per wave in LDS and appended to a device ring CAP at a time.
// Every wave of the chip adds to the ring's one counter; with one atomic per
// window that holds a hit, a hit-dense search (tens of hits per prime) kept
// its waves queued on that address and walked at about 0.8 of its issue
// rate.  Staged, a wave takes one atomic per CAP hits.  Slots past the
// ring's capacity are counted but not written (the host re-runs then).
// `stage` and `fill` are this wave's LDS records and staged count; put()
// and flush() are called by all the wave's active lanes together, and the
// active set differs from call to call (a chunk is re-walked by the lanes
// whose pre-test fired), so the count lives in LDS, where every lane sees
// the last update: `fill` is volatile, or the compiler would keep a copy of
// it per lane, and lanes left out of a call would keep a stale one.
template <class Ring, uint32_t CAP>
struct StagedSink {
  Ring ring;
  typename Ring::Rec *stage;
  volatile uint32_t *fill;

  __device__ __forceinline__ void flush_n(uint32_t f) const {
    const uint64_t act = __ballot(true);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)act) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader) base = atomicAdd(ring.count, (unsigned long long)f);
    base = __shfl(base, leader);
    // the active lanes copy the f records between them, in rank order
    const uint32_t nact = (uint32_t)__popcll(act);
    const uint32_t rank = (uint32_t)__popcll(act & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    for (uint32_t i = rank; i < f; i += nact)
      if (base + i < ring.cap) ring.hits[base + i] = stage[i];
  }
  __device__ __forceinline__ void put(bool hit, uint64_t p, uint64_t k, uint32_t n, int32_t c) const {
    const uint64_t mask = __ballot(hit);
    if (mask == 0) return;
    const uint32_t nh = (uint32_t)__popcll(mask);
    uint32_t f = *fill;
    if (f + nh > CAP) {
      flush_n(f);
      f = 0;
    }
    if (hit) {
      const uint32_t lane = __lane_id();
      const uint64_t below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
      stage[f + (uint32_t)__popcll(below)] = ring.make(p, k, n, c);
    }
    *fill = f + nh;
  }
  __device__ __forceinline__ void flush() const {
    const uint32_t f = *fill;
    if (f) {
      flush_n(f);
      *fill = 0;
    }
  }
};

// Exact window test (pps/app.c:1093-1109 semantics) on a state reduced to
// [0, p): K = k*2^i with i < s, kmin <= k <= kmax, n+i inside [nmin, nmax).
__device__ __forceinline__ bool exact(uint64_t K, uint32_t n, const WindowRule &a, uint64_t &k, uint32_t &nn) {
  const uint32_t lo = lo32(K);
  if (lo == 0) return false;
  const uint32_t i = (uint32_t)__builtin_ctz(lo);
  if (i >= a.s) return false;
  k = K >> i;
  nn = n + i;
  return k <= a.kmax && k >= a.kmin && nn < a.nmax;
}

// Pre-test of one window on the canonical state K: an odd part <= kmax for
// the walked sign (both signs: also for p - K); pmb = p - bound.
template <int SIGN, int PRE>
__device__ __forceinline__ bool pre(uint64_t K, uint64_t p, uint64_t pmb, const WindowRule &a) {
  if (PRE == PRE_RANGE) {
    if (SIGN == 0) return (int)(K <= a.bound) | (int)(K >= pmb);
    return K <= a.bound;
  }
  if (SIGN == 0) return (int)pretest(K, a.kmax) | (int)pretest(p - K, a.kmax);
  return pretest(K, a.kmax);
}

// Per-lane constants of the R primes a lane walks, and a snapshot of their
// states; passed by value so the rare re-walk call needs no scratch memory.
template <int R>
struct LanePrimes {
  uint64_t p[R];
  uint64_t pmb[R];  // p - bound (range pre-test, both signs)
  uint32_t ps[R];
  bool valid[R];
  uint32_t zero;    // opaque_zero(), for step()
};
template <int R>
struct LaneStates {
  uint64_t K[R];
};

// Re-walk `cnt` windows from state K (window start n), for chunks whose
// states were not kept: the same per-window pre-test, full test and
// emission as recheck, stepping again between windows.  FUSED: the wave's
// step (step_fused), as in test_and_step.
template <int SIGN, int R, bool S32, bool FUSED, int RPRE, class Sink>
__device__ __forceinline__ void rescan(const LanePrimes<R> L, LaneStates<R> S, uint32_t n, uint32_t cnt,
                                       const WindowRule &a, const Sink &sink) {
  for (uint32_t u = 0; u < cnt; ++u, n += a.s) {
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; ++r) any |= pre<SIGN, RPRE>(S.K[r], L.p[r], L.pmb[r], a);
    if (__ballot(any)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t Kc = S.K[r];
        uint64_t k = 0;
        uint32_t nn = 0;
        if (SIGN == 0) {
          bool h = L.valid[r] && exact(Kc, n, a, k, nn);  // Kc is 2^-n: the -1 form
          sink.put(h, L.p[r], k, nn, -1);
          h = L.valid[r] && exact(L.p[r] - Kc, n, a, k, nn);
          sink.put(h, L.p[r], k, nn, +1);
        } else {
          const bool h = L.valid[r] && exact(Kc, n, a, k, nn);
          sink.put(h, L.p[r], k, nn, SIGN);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      S.K[r] = FUSED ? step_fused<S32>(S.K[r], L.p[r], L.ps[r], a.s, L.zero)
                     : step<S32>(S.K[r], L.p[r], L.ps[r], a.s, L.zero);
  }
}

// Re-check the W windows of a chunk whose pre-test fired, from the states
// the chunk passed through (kept in registers while it was walked, so no
// step is repeated).  Reached only when a pre-test fired somewhere in the
// chunk: a false alarm of a cheap form, or a hit.  Each window first gets a
// per-window pre-test, RPRE: the exact one (three instructions per state),
// or for the range form its 64-bit compare (one per sign, which is all the
// alarm needs to be located); only a window that passes it (wave-wide) gets
// the full test and the emission.
template <int SIGN, int R, int W, int RPRE, class Sink>
__device__ __forceinline__ void recheck(const LanePrimes<R> &L, const LaneStates<R> (&hist)[W], uint32_t n,
                                        const WindowRule &a, const Sink &sink) {
#pragma unroll
  for (int u = 0; u < W; ++u, n += a.s) {
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; ++r) any |= pre<SIGN, RPRE>(hist[u].K[r], L.p[r], L.pmb[r], a);
    if (__ballot(any)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t Kc = hist[u].K[r];
        uint64_t k = 0;
        uint32_t nn = 0;
        if (SIGN == 0) {
          bool h = L.valid[r] && exact(Kc, n, a, k, nn);  // Kc is 2^-n: the -1 form
          sink.put(h, L.p[r], k, nn, -1);
          h = L.valid[r] && exact(L.p[r] - Kc, n, a, k, nn);
          sink.put(h, L.p[r], k, nn, +1);
        } else {
          const bool h = L.valid[r] && exact(Kc, n, a, k, nn);
          sink.put(h, L.p[r], k, nn, SIGN);
        }
      }
    }
  }
}

// Pre-test of W windows of R states each, stepping them and keeping each
// window's states in hist: true when one of them may hold a hit.  FUSED:
// the three-multiply step, for a wave whose primes all pass fused_step_ok.
template <int SIGN, int R, int W, bool S32, int PRE, bool FUSED>
__device__ __forceinline__ bool test_and_step(const LanePrimes<R> &L, LaneStates<R> &S, const WindowRule &a,
                                              LaneStates<R> (&hist)[W]) {
  bool any = false;
  uint32_t lowest = ~0u;
  uint32_t highest[R];
#pragma unroll
  for (int r = 0; r < R; ++r) highest[r] = 0;
#pragma unroll
  for (int u = 0; u < W; ++u) {
    hist[u] = S;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (PRE == PRE_ODDLO) {
        lowest = min(lowest, odd_low(S.K[r]));
        if (SIGN == 0) lowest = min(lowest, odd_low(L.p[r] - S.K[r]));
      } else if (PRE == PRE_RANGE) {
        // K <= bound needs hi32(K) <= hi32(bound), and K >= p - bound needs
        // hi32(K) >= hi32(p - bound): a running minimum (and, both signs, a
        // maximum per prime) of the high word, v_min3_u32 / v_max3_u32 per
        // two windows, in place of a 64-bit compare per window and sign
        lowest = min(lowest, hi32(S.K[r]));
        if (SIGN == 0) highest[r] = max(highest[r], hi32(S.K[r]));
      } else {
        any |= pre<SIGN, PRE>(S.K[r], L.p[r], L.pmb[r], a);
      }
      S.K[r] = FUSED ? step_fused<S32>(S.K[r], L.p[r], L.ps[r], a.s, L.zero)
                     : step<S32>(S.K[r], L.p[r], L.ps[r], a.s, L.zero);
    }
  }
  if (PRE == PRE_ODDLO) any = lowest <= (uint32_t)a.kmax;
  if (PRE == PRE_RANGE) {
    any = lowest <= hi32(a.bound);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (SIGN == 0) any |= highest[r] >= hi32(L.pmb[r]);
  }
  return any;
}

// A chunk whose pre-test fired is checked window by window: with KEEP from
// the states it passed through (U*R more state registers, no re-stepping),
// else by re-walking it from a snapshot.  KEEP pays where false alarms are
// frequent (the range form, on up to a tenth of the chunks) and where the
// chunk is short enough that its U states cost no resident waves.
template <int SIGN, int R, int U, bool S32, int PRE, bool FUSED, bool KEEP, class Sink>
__device__ __forceinline__ void walk_chunks(const LanePrimes<R> &L, LaneStates<R> &S, uint32_t n, uint32_t nwin,
                                            const WindowRule &a, const Sink &sink) {
  constexpr int RPRE = PRE == PRE_RANGE ? PRE_RANGE : PRE_EXACT;
  const uint32_t s = a.s;
  uint32_t w = 0;
  // One chunk: its pre-test while stepping, then, if it fired, its windows
  // checked one by one -- from the states kept in hist (KEEP) or by walking
  // the chunk again from its snapshot.
  auto chunk = [&](uint32_t n0) {
    LaneStates<R> hist[U];
    const LaneStates<R> snap = S;
    if (test_and_step<SIGN, R, U, S32, PRE, FUSED>(L, S, a, hist)) {
      if (KEEP)
        recheck<SIGN, R, U, RPRE>(L, hist, n0, a, sink);
      else
        rescan<SIGN, R, S32, FUSED, RPRE>(L, snap, n0, U, a, sink);
    }
  };
  // Two chunks a turn: a chunk's first state must stay in its registers
  // through the chunk (it is the snapshot, or the first kept state), so with
  // one chunk a turn the loop ended by copying the last state into them (a
  // v_mov_b64 per chunk, which takes an issue slot); with two, the state
  // alternates between two register pairs and no copy is left.
  for (; w + 2 * U <= nwin; w += 2 * U) {
    chunk(n);
    n += U * s;
    chunk(n);
    n += U * s;
  }
  if (w + U <= nwin) {
    chunk(n);
    n += U * s;
    w += U;
  }
  for (; w < nwin; ++w, n += s) {
    LaneStates<R> hist[1];
    if (test_and_step<SIGN, R, 1, S32, PRE, FUSED>(L, S, a, hist)) recheck<SIGN, R, 1, RPRE>(L, hist, n, a, sink);
  }
}

// Walk nwin windows from window start n, states S.  Steps run in chunks of U
// and a chunk only leaves the straight-line path when its pre-test fired
// (then it is walked again with exact tests).  A wave whose primes all pass
// fused_step_ok walks with the three-multiply step (a wave-uniform choice,
// so no lane diverges).
template <int SIGN, int R, int U, bool S32, int PRE, bool KEEP, class Sink>
__device__ __forceinline__ void walk_windows(const LanePrimes<R> &L, LaneStates<R> &S, uint32_t n, uint32_t nwin,
                                             const WindowRule &a, const Sink &sink) {
  bool ok = true;
#pragma unroll
  for (int r = 0; r < R; ++r) ok = ok && fused_step_ok(L.p[r], a.s);
  if (__all(ok))
    walk_chunks<SIGN, R, U, S32, PRE, true, KEEP>(L, S, n, nwin, a, sink);
  else
    walk_chunks<SIGN, R, U, S32, PRE, false, KEEP>(L, S, n, nwin, a, sink);
}

// Start states: K = -c * 2^-e mod p (both signs: 2^-e, whose negation p - K
// is the +1 form).
template <int SIGN, int R>
__device__ __forceinline__ void start_states(LanePrimes<R> &L, LaneStates<R> &S, uint32_t e, uint32_t s,
                                             uint64_t bound) {
  L.zero = opaque_zero();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    L.pmb[r] = L.p[r] - bound;
    const uint64_t pinv = inv64(L.p[r]);
    L.ps[r] = step_multiplier(pinv, s);
    const uint64_t x = inv2pow(L.p[r], pinv, e);
    S.K[r] = (SIGN == 1) ? L.p[r] - x : x;
  }
}

}  // namespace ppsd
