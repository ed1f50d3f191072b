// ppsd_arith.h -- exact 64-bit modular arithmetic for the PPSieve n-walk on
// gfx950 (wave64).  Integer only; no floating point reaches a residue.
//
// Notation: p odd, 2^32 < p < 2^62.  pinv = p^-1 mod 2^64 (positive inverse).
// REDC is used in its "difference" form: for T < p*2^64 and m = lo64(T)*pinv,
// lo64(m*p) == lo64(T), so T*2^-64 == hi64(T) - hi64(m*p) (mod p) exactly,
// and the difference lies in (-p, p): one conditional +p makes it canonical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppsd {

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

// p^-1 mod 2^64 by Newton's iteration x <- x(2 - p x), starting from x = p,
// which is correct to 3 bits because p^2 == 1 (mod 8): 3, 6, 12, 24, 48 bits
// in 32-bit arithmetic, then one 64-bit step.
__device__ __forceinline__ uint64_t inv64(uint64_t p) {
  const uint32_t p32 = lo32(p);
  uint32_t x = p32;
  x *= 2u - p32 * x;
  x *= 2u - p32 * x;
  x *= 2u - p32 * x;
  x *= 2u - p32 * x;
  uint64_t y = x;
  y *= 2u - p * y;
  return y;
}

// a*b*2^-64 mod p for a, b < p (canonical result).
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t pinv) {
  const uint64_t lo = a * b;
  const uint64_t hi = __umul64hi(a, b);
  const uint64_t m = lo * pinv;
  const uint64_t mh = __umul64hi(m, p);
  uint64_t t;
  const bool neg = __builtin_sub_overflow(hi, mh, &t);
  return neg ? t + p : t;
}

// x/2 mod p for canonical x.
__device__ __forceinline__ uint64_t half_mod(uint64_t x, uint64_t p) { return (x & 1) ? (x + p) >> 1 : x >> 1; }

// 2^-e mod p (canonical) for e >= 32 and p > 2^32.  The Montgomery form of
// 2^-t is 2^(64-t); starting from t = the top six bits of e (32 <= t <= 63)
// it is already reduced, and the remaining bits of e are consumed left to
// right: square, then halve where the bit is set.  A final REDC leaves
// Montgomery form.
__device__ __forceinline__ uint64_t inv2pow(uint64_t p, uint64_t pinv, uint32_t e) {
  const int L = 32 - __clz((int)e);  // bit length of e, >= 6
  const uint32_t t = e >> (L - 6);
  uint64_t x = (uint64_t)1 << (64 - t);
  for (int b = L - 7; b >= 0; --b) {
    x = mont_mul(x, x, p, pinv);
    if ((e >> b) & 1u) x = half_mod(x, p);
  }
  // REDC(x) = x * 2^-64: T = x, hi(T) = 0
  const uint64_t m = x * pinv;
  const uint64_t mh = __umul64hi(m, p);
  return mh ? p - mh : 0;
}

// One window: K <- K * 2^-s (mod p) for 1 <= s <= 32, as the exact division
// K' = (K + j*p) / 2^s with j = K * (-p^-1) mod 2^s (so K + j*p == 0 mod 2^s).
// Splitting off the low s bits, K' = (K >> s) + ceil(j*p / 2^s), and with
// m = j * 2^(32-s):  ceil(j*p / 2^s) = m*p_hi + ceil(m*p_lo / 2^32), where
// ceil(m*p_lo / 2^32) is the high word of m*p_lo + (2^32 - 1).  Folding the
// 2^(32-s) into the inverse, m = K_lo * ps_sh mod 2^32 with
// ps_sh = (-p^-1 mod 2^32) << (32-s).
// Every term is non-negative, so there is no borrow to repair, and a
// canonical state stays canonical: K <= p-1 and j <= 2^s - 1 give
// K + j*p <= 2^s*p - 1, so K' <= p - 1.  No per-step reduction, and the
// walk's pre-test never sees an unreduced state.
//
// The last addition, of the 32-bit ceiling term to a 64-bit value, is a
// v_mad_u64_u32 with multiplier 1 (add_u32_u64): the term sits in the high
// register of the multiply that made it, and an ordinary 64-bit add would
// first move it next to a zero register.
__device__ __forceinline__ uint64_t add_u32_u64(uint32_t a, uint64_t b) {
  uint64_t r, carry;
  asm("v_mad_u64_u32 %0, %1, %2, 1, %3" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
  return r;
}

//
// At s = 32 the middle mad's 64-bit addend is hi32(K) zero-extended, an
// even-aligned register pair whose high register must hold 0.  `zero` is 0
// from opaque_zero(): the compiler cannot re-create it with a move, so it
// keeps it in one register and each window costs a single v_mov_b32 (of
// hi32(K) into the register below it).  With a literal 0 the register
// allocator sometimes rewrote the zero every window as well.
__device__ __forceinline__ uint32_t opaque_zero() {
  uint32_t z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// a*b + c as one v_mad_u64_u32.  Written as asm: with a second, fused walk
// loop beside it (step_fused) the compiler widened the plain expression's
// product to 64 x 32 bits, two mads per product.
__device__ __forceinline__ uint64_t mad_u64_u32(uint32_t a, uint32_t b, uint64_t c) {
  uint64_t r, carry;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(c));
  return r;
}

template <bool S32>
__device__ __forceinline__ uint64_t step(uint64_t K, uint64_t p, uint32_t ps_sh, uint32_t s, uint32_t zero) {
  const uint32_t m = lo32(K) * ps_sh;
  const uint64_t ks = S32 ? ((uint64_t)zero << 32 | hi32(K)) : (K >> s);
  const uint64_t u = mad_u64_u32(m, hi32(p), ks);
  const uint32_t hc = hi32(mad_u64_u32(m, lo32(p), 0xffffffffull));
  return add_u32_u64(hc, u);
}

// The window step in three multiplies, for primes with
// ((p-1) >> s) + lo32(p) < 2^32 (at s = 32: lo32(p) + hi32(p) < 2^32).
// With L = K << (32-s) mod 2^64 -- the 64-bit value X*2^32 + (lo_s(K) <<
// (32-s)) for X = K >> s, which needs X < 2^32 -- the sum
// S = m*p_lo + L has a zero low word (lo32(m*p_lo) == -(lo_s(K) << (32-s))
// mod 2^32) and the high word X + hi32(m*p_lo) + [lo_s(K) != 0], which is
// X + ceil(m*p_lo / 2^32); so K' = m*p_hi + hi32(S).  S < 2^64 because
// m*p_lo <= (2^32-1)*p_lo and L < (X+1)*2^32 with X <= (p-1) >> s, which the
// bound above keeps below 2^64 in sum.  At s = 32, L is K itself: the first
// mad takes the state's own register pair as its addend, with no move.  The
// step is then one v_mul_lo_u32 and two v_mad_u64_u32 (plus v_lshlrev_b64 at
// s < 32, where step() needs a 64-bit shift too), and one v_mov_b32 that
// places hi32(S) beside the zero register for the second mad.  Primes that
// fail the bound have lo32(p) within (p-1) >> s of 2^32, i.e. they lie at
// the top of each 2^32 block, so in an ascending batch they form runs and
// whole waves take one step or the other.
__device__ __forceinline__ bool fused_step_ok(uint64_t p, uint32_t s) {
  return ((p - 1) >> s) + lo32(p) <= 0xffffffffull;
}
template <bool S32>
__device__ __forceinline__ uint64_t step_fused(uint64_t K, uint64_t p, uint32_t ps_sh, uint32_t s, uint32_t zero) {
  const uint32_t m = lo32(K) * ps_sh;
  const uint64_t L = S32 ? K : K << (32u - s);
  const uint32_t hs = hi32(mad_u64_u32(m, lo32(p), L));
  return mad_u64_u32(m, hi32(p), (uint64_t)zero << 32 | hs);
}

// The multiplier of step() for prime p and window s.
__device__ __forceinline__ uint32_t step_multiplier(uint64_t pinv, uint32_t s) {
  return (0u - lo32(pinv)) << (32u - s);
}

// Find-first-set-bit of the low word, as v_ffbl_b32: 0xffffffff when the word
// is zero.  Written as asm so a zero word is defined behaviour (no hit is
// possible then, because the trailing-zero count is >= 32 >= s; the shifted
// value only feeds a pre-test that the exact re-check repeats).
__device__ __forceinline__ uint32_t ffbl(uint32_t x) {
  uint32_t r;
  asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Necessary condition for a hit in the window whose (canonical) state is K:
// the odd part of K must not exceed kmax.
__device__ __forceinline__ bool pretest(uint64_t K, uint64_t kmax) {
  const uint32_t i = ffbl(lo32(K));
  return (K >> (i & 63u)) <= kmax;
}

}  // namespace ppsd
