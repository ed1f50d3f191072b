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

// Negative powers of two modulo each small prime q of the filter's table,
// u16 each, kPowPer per prime in the table's order: 2^-j mod q for j <
// kPowLo, then 2^(-kPowLo*j) mod q for j < kPowHi.  Since e = n mod (q-1) <
// 2^15 = kPowLo*kPowHi, 2^-e mod q is the product of one entry of each half.
constexpr uint32_t kPowLoBits = 7;
constexpr uint32_t kPowLo = 1u << kPowLoBits, kPowHi = 256, kPowPer = kPowLo + kPowHi;
constexpr uint32_t kPowPassBytes = 64 * kPowPer * sizeof(uint16_t);  // one pass of 64 primes: 48 KiB

// q | k*2^n + c ?  (q an odd prime below 2^15, c = +1, or -1 for NEG.)
// Since 2 is a unit mod q, that is q | k + c*2^-n, and 2^-n = 2^-e with
// e = n mod (q-1) (Fermat) = lo*hi mod q for one entry of each half of q's
// row (in LDS).  kq is k itself when k < 2^31, else k mod q from both
// 32-bit halves (the high half times 2^32 mod q).  x = kq + lo*hi (c = +1)
// or kq + lo*(q - hi) (c = -1, since q - hi == -hi) is congruent to
// k + c*2^-n and below 2^31 + 2^30, so the answer is whether q | x: one
// multiply by q^-1 mod 2^32 and one compare.  About 17 VALU instructions
// per (hit, q).
__device__ __forceinline__ uint32_t k_mod_q(uint32_t klo, uint32_t khi, const SmallPrime &sp) {
  const uint32_t a = mod_barrett(khi, sp.q, sp.mq) * sp.c32 + mod_barrett(klo, sp.q, sp.mq);  // < 2^31
  return mod_barrett(a, sp.q, sp.mq);
}
template <bool NEG>
__device__ __forceinline__ bool divides_tab(uint32_t kq, uint32_t n, const SmallPrime &sp, const uint16_t *row) {
  const uint32_t e = mod_barrett(n, sp.q - 1, sp.md);
  const uint32_t lo = row[e & (kPowLo - 1)], hi = row[kPowLo + (e >> kPowLoBits)];
  const uint32_t x = kq + lo * (NEG ? sp.q - hi : hi);
  return x * sp.qinv <= sp.qlim;
}

// Of the live hits in `mask` (wave-uniform; those in `big` have k >= 2^31),
// the ones the lane's prime q divides, for one sign: each hit is read into
// scalar registers and tested against the wave's 64 primes.  Hits with
// k < 2^31 go kHitsPerTurn at a time (the last turn repeats its last hit to
// fill up), so that independent chains of multiplies and LDS reads
// interleave.
constexpr int kHitsPerTurn = 2;
template <bool NEG>
__device__ __forceinline__ uint64_t divisible_hits(uint64_t mask, uint64_t big, uint64_t k, uint32_t n,
                                                   const SmallPrime &sp, const uint16_t *row) {
  uint64_t dead = 0;
  for (uint64_t m = mask & ~big; m;) {
    int j[kHitsPerTurn];
#pragma unroll
    for (int h = 0; h < kHitsPerTurn; ++h) {
      j[h] = h ? j[h - 1] : 0;
      if (m) {
        j[h] = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
      }
    }
    bool d[kHitsPerTurn];
#pragma unroll
    for (int h = 0; h < kHitsPerTurn; ++h) {
      const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, j[h]);
      const uint32_t nj = (uint32_t)__builtin_amdgcn_readlane((int)n, j[h]);
      d[h] = divides_tab<NEG>(kj, nj, sp, row);
    }
#pragma unroll
    for (int h = 0; h < kHitsPerTurn; ++h)
      if (__ballot(d[h])) dead |= 1ull << j[h];
  }
  for (uint64_t m = mask & big; m;) {
    const int j = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    const uint32_t klo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, j);
    const uint32_t khi = (uint32_t)__builtin_amdgcn_readlane((int)(k >> 32), j);
    const uint32_t nj = (uint32_t)__builtin_amdgcn_readlane((int)n, j);
    if (__ballot(divides_tab<NEG>(k_mod_q(klo, khi, sp), nj, sp, row))) dead |= 1ull << j;
  }
  return dead;
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

// Workgroup shape of the filter kernels: kFilterThreads threads, each
// holding kFilterSlots hits, and one pass of powers of two in LDS (48 KiB:
// three workgroups fit a CU's 160 KiB, 24 waves).
constexpr uint32_t kFilterThreads = 512;
constexpr int kFilterSlots = 4;
constexpr uint32_t kFilterGrid = 768;  // workgroups per launch: three per CU of MI355X's 256

// Filters raw hits [0, nraw) into a kept ring.  A workgroup takes
// kFilterThreads * kFilterSlots hits at a time (kFilterSlots per lane) and
// applies the cheap part of the rule lane-parallel.  try_all_factors then
// runs pass by pass over the small primes, 64 per pass, smallest first: the
// workgroup stages the pass's powers of two (pow2, kPowPer per prime) in LDS
// (lds: kPowPassBytes, 16-byte aligned), and each wave tests its surviving
// hits one at a time against the pass's 64 primes, one prime per lane.  A hit
// with a divisor drops out; most composites do in the first passes, and the
// workgroup stops when none of its hits is left.  A kept hit is written by
// the lane that holds it.  rule(x, kstep, koffset, nmin, nmax) gives the
// hit's rule; nsmall is a multiple of 64.  Every thread of the workgroup
// must call this (it synchronises the workgroup).
template <class HitT, class Rule>
__device__ __forceinline__ void filter_hits(const HitT *hits, uint64_t nraw, HitT *kept,
                                            unsigned long long *kept_count, uint64_t kept_cap,
                                            const SmallPrime *small_primes, const uint16_t *pow2, uint32_t nsmall,
                                            const Rule &rule, uint16_t *lds) {
  const uint32_t lane = __lane_id();
  const uint64_t per = (uint64_t)blockDim.x * kFilterSlots;
  const uint64_t first = (uint64_t)(threadIdx.x >> 6) * 64 * kFilterSlots + lane;
  const uint16_t *row = lds + lane * kPowPer;  // this lane's prime's powers, in every pass
  for (uint64_t base = blockIdx.x * per; base < nraw; base += (uint64_t)gridDim.x * per) {  // workgroup-uniform
    HitT x[kFilterSlots];
    bool alive[kFilterSlots];
#pragma unroll
    for (int s = 0; s < kFilterSlots; ++s) {
      const uint64_t h = base + first + 64 * s;
      const bool valid = h < nraw;
      x[s] = HitT{};
      if (valid) x[s] = hits[h];
      uint32_t kstep = 2, koffset = 1, nmin = 0, nmax = 0;
      if (valid) rule(x[s], kstep, koffset, nmin, nmax);
      alive[s] = valid && rule_cheap(x[s].k, x[s].n, x[s].c, kstep, koffset, nmin, nmax);
    }
    for (uint32_t j0 = 0; j0 < nsmall; j0 += 64) {
      __syncthreads();  // the previous pass's table is no longer read
      const uint4 *src = (const uint4 *)(pow2 + (size_t)j0 * kPowPer);
      uint4 *dst = (uint4 *)lds;
      for (uint32_t i = threadIdx.x; i < kPowPassBytes / 16; i += blockDim.x) dst[i] = src[i];
      const SmallPrime sp = small_primes[j0 + lane];
      bool any = false;
#pragma unroll
      for (int s = 0; s < kFilterSlots; ++s) any |= alive[s];
      if (!__syncthreads_or(any)) break;  // also the barrier after staging
#pragma unroll
      for (int s = 0; s < kFilterSlots; ++s) {
        // wave-uniform masks of the slot: live hits, those with k >= 2^31,
        // those with c = -1 (ballots once per pass, so that only the live
        // mask's walk holds scalar registers through the hit loop)
        const uint64_t live = __ballot(alive[s]);
        if (!live) continue;
        const uint64_t big = __ballot(x[s].k >= 0x80000000ull);
        const uint64_t neg = __ballot(x[s].c < 0);
        const uint64_t dead = divisible_hits<false>(live & ~neg, big, x[s].k, x[s].n, sp, row) |
                              divisible_hits<true>(live & neg, big, x[s].k, x[s].n, sp, row);
        if ((dead >> lane) & 1) alive[s] = false;
      }
    }
#pragma unroll
    for (int s = 0; s < kFilterSlots; ++s) {
      const uint64_t keep = __ballot(alive[s]);
      const unsigned long long idx = wave_slot(keep, kept_count);
      if (alive[s] && idx < kept_cap) kept[idx] = x[s];
    }
  }
}

}  // namespace ppsd
