// ppsd_filter.h -- test_factor (pps/app.c:159-186) on the device: the k
// residue rule, the n range, the shifted-k residue pre-tests with their
// k-size guards, and try_all_factors (pps/factor_proth.c:220-253).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppsd_kernels.h"

namespace ppsd {

// x mod d for any 32-bit x, given m = floor(2^32 / d) and 1 < d < 2^31: the
// estimate hi32(x*m) is the quotient or one less (m > 2^32/d - 1, x < 2^32),
// so x - estimate*d lies in [0, 2d) and one unsigned min finishes it.
__device__ __forceinline__ uint32_t mod_barrett(uint32_t x, uint32_t d, uint32_t m) {
  const uint32_t r = x - __umulhi(x, m) * d;
  return min(r, r - d);
}

// q | k*2^n + c ?  (k < 2^62, q an odd prime below 2^15)  k mod q from its
// two 32-bit halves; 2^n mod q as 2^e with e = n mod (q-1) (Fermat), its top
// four bits at once -- 2^(e>>11) < 2q since e < q-1 -- then eleven
// square-and-maybe-double steps, each product below 2q^2 < 2^31.
__device__ __forceinline__ bool divides(uint64_t k, uint32_t n, int32_t c, const SmallPrime &sp) {
  const uint32_t q = sp.q;
  const uint32_t kl = mod_barrett((uint32_t)k, q, sp.mq);
  const uint32_t kh = mod_barrett((uint32_t)(k >> 32), q, sp.mq);
  const uint32_t km = mod_barrett(kh * sp.c32 + kl, q, sp.mq);
  const uint32_t e = mod_barrett(n, q - 1, sp.md);
  uint32_t t = 1u << (e >> 11);
  t = min(t, t - q);
#pragma unroll
  for (int b = 10; b >= 0; --b) t = mod_barrett((t * t) << ((e >> b) & 1u), q, sp.mq);
  const uint32_t v = mod_barrett(km * t, q, sp.mq);
  return v == (c > 0 ? q - 1 : 1u);
}

// The residue pre-tests of test_factor in wrapping 64-bit arithmetic with their
// k-size guards; mod31 receives (k << (n % 10)) + c when that branch is
// reached and stays 1 otherwise, as in the reference.
__device__ __forceinline__ bool pre_rules(uint64_t k, uint32_t n, int32_t c, uint64_t &mod31) {
  const uint64_t cc = (uint64_t)(int64_t)c;
  const uint32_t khigh = (uint32_t)(k >> 32);
  const uint32_t r15 = (uint32_t)(((k << (n & 3)) + cc) % 15);
  // 1 where r15 is divisible by neither 3 nor 5 (pps/app.c:152)
  const bool ok15 = (0x6996u >> r15) & 1u;
  mod31 = 1;
  bool keep = ok15 && (uint32_t)(((k << (n % 3)) + cc) % 7) != 0;
  if (keep && khigh < (1u << 24)) {
    keep = (uint32_t)(((k << (n & 7)) + cc) % 17) != 0;
    if (keep && khigh < (1u << 22)) {
      mod31 = (k << (n % 10)) + cc;
      keep = (uint32_t)(mod31 % 11) != 0;
      if (keep && khigh < (1u << 21)) {
        keep = (uint32_t)(((k << (n % 11)) + cc) % 23) != 0;
        if (keep && khigh < (1u << 20)) {
          keep = (uint32_t)(((k << (n % 12)) + cc) % 13) != 0;
          if (keep && khigh < (1u << 14)) keep = (uint32_t)(((k << (n % 18)) + cc) % 19) != 0;
        }
      }
    }
  }
  return keep;
}

// The rule's cheap part for one hit, per lane: k in the residue class, n in
// range, the residue pre-tests.
__device__ __forceinline__ bool rule_cheap(uint64_t k, uint32_t n, int32_t c, uint32_t kstep, uint32_t koffset,
                                           uint32_t nmin, uint32_t nmax) {
  const uint64_t b = k / kstep;
  bool keep = k == (uint64_t)kstep * b + koffset && n >= nmin && n < nmax;
  uint64_t mod31 = 1;
  if (keep) keep = pre_rules(k, n, c, mod31) && (uint32_t)(mod31 % 31) != 0;
  return keep;
}

// try_all_factors for one (k, n, c) held by every lane of the wave: the
// lanes split the small primes, 64 per pass, smallest first; most composites
// fall in the first passes, so the wave stops at the first pass that finds a
// divisor.  Each pass loads the next pass's constants before it computes.
// nsmall is a multiple of 64.  Returns the same value in every lane.
__device__ __forceinline__ bool no_small_factor(uint64_t k, uint32_t n, int32_t c, const SmallPrime *small_primes,
                                                uint32_t nsmall) {
  const uint32_t lane = __lane_id();
  SmallPrime cur = small_primes[lane];
  for (uint32_t j0 = 0; j0 < nsmall; j0 += 64) {
    const SmallPrime next = small_primes[j0 + 64 < nsmall ? j0 + 64 + lane : lane];
    if (__ballot(divides(k, n, c, cur)) != 0) return false;
    cur = next;
  }
  return true;
}

// Slots for the lanes set in keep (the same mask in every lane) at the end
// of a ring, in lane order, with one atomic per wave: a counter every kept
// hit incremented on its own serialised the filter on hit-dense units.
// Returns this lane's slot (meaningful only where its bit is set).
__device__ __forceinline__ unsigned long long wave_slot(uint64_t keep, unsigned long long *count) {
  const uint32_t lane = __lane_id();
  unsigned long long base = 0;
  if (keep) {
    if (lane == (uint32_t)(__ffsll((unsigned long long)keep) - 1)) base = atomicAdd(count, (unsigned long long)__popcll(keep));
    base = __shfl(base, __ffsll((unsigned long long)keep) - 1);
  }
  return base + (unsigned long long)__popcll(keep & ((1ull << lane) - 1));
}

// Filters raw hits [0, nraw) into a kept ring.  Each wave takes 64 hits at a
// time, one per lane, applies the cheap part of the rule lane-parallel, then
// runs the trial division wave-wide for each survivor in turn; a kept hit is
// written by the lane that holds it.  rule(x, kstep, koffset, nmin, nmax)
// gives the hit's rule.
template <class HitT, class Rule>
__device__ __forceinline__ void filter_hits(const HitT *hits, uint64_t nraw, HitT *kept,
                                            unsigned long long *kept_count, uint64_t kept_cap,
                                            const SmallPrime *small_primes, uint32_t nsmall, const Rule &rule) {
  const uint32_t lane = __lane_id();
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wave * 64; base < nraw; base += nwaves * 64) {
    const uint64_t h = base + lane;
    const bool valid = h < nraw;
    HitT x{};
    if (valid) x = hits[h];
    uint32_t kstep = 2, koffset = 1, nmin = 0, nmax = 0;
    if (valid) rule(x, kstep, koffset, nmin, nmax);
    const bool cand = valid && rule_cheap(x.k, x.n, x.c, kstep, koffset, nmin, nmax);
    uint64_t mask = __ballot(cand);
    uint64_t keepmask = 0;  // wave-uniform: each verdict below is the wave's
    while (mask) {
      const int j = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      // the hit of lane j, read into scalar registers (j is wave-uniform)
      const uint64_t k = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x.k >> 32), j) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x.k, j);
      const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)x.n, j);
      const int32_t c = __builtin_amdgcn_readlane((int)x.c, j);
      if (no_small_factor(k, n, c, small_primes, nsmall)) keepmask |= 1ull << j;
    }
    const unsigned long long idx = wave_slot(keepmask, kept_count);
    if (((keepmask >> lane) & 1) && idx < kept_cap) kept[idx] = x;
  }
}

}  // namespace ppsd
