// ppsd_filter.h -- test_factor (pps/app.c:159-186) on the device: the k
// residue rule, the n range, the shifted-k residue pre-tests with their
// k-size guards, and try_all_factors (pps/factor_proth.c:220-253).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppsd {

// x*y mod q for x, y < q < 2^15: the float quotient estimate is off by at most
// one (relative error ~2^-22 on a quotient below 2^15), so two corrections
// make it exact.
__device__ __forceinline__ uint32_t mulmod_q(uint32_t x, uint32_t y, uint32_t q, float rq) {
  const uint32_t prod = x * y;
  const uint32_t e = (uint32_t)((float)prod * rq);
  int32_t r = (int32_t)(prod - e * q);
  r += (r < 0) ? (int32_t)q : 0;
  r -= (r >= (int32_t)q) ? (int32_t)q : 0;
  return (uint32_t)r;
}

// q | k*2^n + c ?  2^n mod q via 2^(n mod (q-1)) (Fermat; q an odd prime),
// so at most 15 squarings.
__device__ __forceinline__ bool divides(uint64_t k, uint32_t n, int32_t c, uint32_t q) {
  const float rq = __frcp_rn((float)q);
  const uint32_t e = n % (q - 1);
  uint32_t t = 1;
#pragma unroll
  for (int b = 14; b >= 0; --b) {
    t = mulmod_q(t, t, q, rq);
    if ((e >> b) & 1u) {
      t <<= 1;
      t -= (t >= q) ? q : 0;
    }
  }
  const uint32_t km = (uint32_t)(k % q);
  const uint32_t v = mulmod_q(km, t, q, rq);
  return (c > 0) ? (v + 1 == q) : (v == 1);
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
// divisor.  Returns the same value in every lane.
__device__ __forceinline__ bool no_small_factor(uint64_t k, uint32_t n, int32_t c, const uint32_t *small_primes,
                                                uint32_t nsmall) {
  for (uint32_t j0 = 0; j0 < nsmall; j0 += 64) {
    const uint32_t j = j0 + __lane_id();
    const bool d = j < nsmall && divides(k, n, c, small_primes[j]);
    if (__ballot(d) != 0) return false;
  }
  return true;
}

// Filters raw hits [0, nraw) into a kept ring.  Each wave takes 64 hits at a
// time, one per lane, applies the cheap part of the rule lane-parallel, then
// runs the trial division wave-wide for each survivor in turn; a kept hit is
// written by the lane that holds it.  rule(x, kstep, koffset, nmin, nmax)
// gives the hit's rule.
template <class HitT, class Rule>
__device__ __forceinline__ void filter_hits(const HitT *hits, uint64_t nraw, HitT *kept,
                                            unsigned long long *kept_count, uint64_t kept_cap,
                                            const uint32_t *small_primes, uint32_t nsmall, const Rule &rule) {
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
    while (mask) {
      const int j = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      const uint64_t k = __shfl(x.k, j);
      const uint32_t n = __shfl(x.n, j);
      const int32_t c = __shfl(x.c, j);
      const bool keep = no_small_factor(k, n, c, small_primes, nsmall);
      if (keep && (int)lane == j) {
        const unsigned long long idx = atomicAdd(kept_count, 1ull);
        if (idx < kept_cap) kept[idx] = x;
      }
    }
  }
}

}  // namespace ppsd
