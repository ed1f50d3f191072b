// ppsd_kernels.h -- kernel argument blocks and launch helpers shared by the
// kernels (ppsd_kernels.hip) and the C-ABI host layer (ppsd_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppsd {

// The constants of one small prime's arithmetic for the factor filter,
// computed on the host.  The device table is padded to whole passes of 64
// with copies of its last prime (a repeated divisor changes no verdict).
struct SmallPrime {
  uint32_t q;
  uint32_t mq;    // floor(2^32 / q)
  uint32_t md;    // floor(2^32 / (q - 1))
  uint32_t c32;   // 2^32 mod q
  uint32_t qinv;  // q^-1 mod 2^32
  uint32_t qlim;  // floor((2^32 - 1) / q): for x < 2^32, q | x iff x * qinv mod 2^32 <= qlim
  uint32_t pad[2];
};

// One raw hit: p | k*2^n + c, with k odd in [kmin,kmax] and nmin <= n < nmax.
// Same field order and widths as ppsd_factor in include/ppsd.h.
struct Hit {
  uint64_t p;
  uint64_t k;
  uint32_t n;
  int32_t c;
};

struct WalkArgs {
  const uint64_t *P;        // primes, any order
  const uint32_t *count_d;  // prime count on the device (range mode) or null
  uint32_t count;           // prime count when count_d is null
  uint32_t cap;             // entries of P: the device count is clamped to it
  uint32_t nmin, nmax;      // nmax exclusive
  uint32_t s;               // window, 1..32
  uint32_t nwin;            // windows to walk: ceil((nmax - nmin) / s)
  uint64_t kmin, kmax;
  int pre;                           // PRE_EXACT, PRE_RANGE or PRE_ODDLO (ppsd_walk.h)
  Hit *hits;                         // raw-hit ring
  unsigned long long *hit_count;     // slots requested (may exceed cap)
  uint64_t hit_cap;
};

struct FilterArgs {
  const Hit *hits;
  const unsigned long long *hit_count;
  uint64_t hit_cap;
  uint64_t kmin, kmax;  // unused by the rule itself, kept for symmetry
  uint32_t nmin, nmax;
  uint32_t kstep, koffset;
  const SmallPrime *small_primes;  // odd primes 11 <= q < 32768, padded
  const uint16_t *pow2;            // their powers of two (ppsd_filter.h kPowPer each)
  uint32_t nsmall;                 // a multiple of 64
  Hit *kept;
  unsigned long long *kept_count;
  uint64_t kept_cap;
  // candidate list of a sieve file (null: the small-prime rule applies):
  // bit (b - cand_b0) of row (n - cand_n0), k = kstep*b + koffset
  const uint8_t *cand;
  uint64_t cand_row_bytes, cand_b0;
  uint32_t cand_n0, cand_rows;
};

struct SieveArgs {
  uint64_t lo;       // odd; bit b <-> lo + 2b
  uint64_t nbits;    // candidates in this chunk
  const uint32_t *q; // sieving primes, ascending
  uint32_t nq_small; // q[0..nq_small) are marked per segment in LDS
  uint32_t nq;       // q[nq_small..nq) are marked with global atomics
  uint32_t *bitmap;  // ceil(nbits/32) words; set bit = composite
  const uint64_t *qinv;  // floor((2^64-1)/q) for q[0..nq_small), or null
  uint32_t block;    // threads per LDS-sieve workgroup: 1024 (default) or 256
};

// Segment geometry of the LDS sieve: 2^18 candidates (2^19 integers) per
// workgroup, 32 KiB of LDS.
constexpr uint32_t kSegBits = 1u << 18;
constexpr uint32_t kSegWords = kSegBits / 32;
constexpr uint32_t kScanBlock = 256;  // words per compaction block
constexpr uint32_t kBlocksPerSeg = kSegWords / kScanBlock;

// One piece of a batched round's sieve: odd candidates lo + 2b, b < nbits,
// held at bits bit0 + b of its launch's bitmap (bit0 a multiple of kSegBits,
// so that each segment belongs to one piece).
struct RoundPiece {
  uint64_t lo, nbits, bit0;
  uint32_t nq_small;  // sieving primes q[0..nq_small) marked in LDS
  uint32_t nq;        // q[nq_small..nq) marked with global atomics
};

// The whole sieve of pieces [p0, p1) of a round in one set of launches,
// whatever their number: the LDS sieve over every segment, the large primes
// over every (piece, prime) pair, one compaction of the launch's bitmap.
struct RoundSieveArgs {
  const RoundPiece *pieces;   // the round's pieces
  uint32_t p0, p1;            // this launch's
  const uint32_t *seg_piece;  // per segment of this launch's bitmap: its piece
  uint32_t nseg;
  const uint32_t *by_nq;      // the launch's pieces in descending order of nq
  uint32_t qlo, qhi;          // large primes of the launch: q[qlo..qhi), qhi = its largest nq
  const uint32_t *q;          // sieving primes, ascending
  const uint64_t *qinv;       // floor((2^64-1)/q) of the LDS-sieved ones
  uint32_t *bitmap;           // nseg * kSegWords words; set bit = composite
  uint32_t block;             // threads per LDS-sieve workgroup: 256 or 1024
  uint32_t grid;              // at most this many workgroups per sieve launch (0: no cap)
};

hipError_t launch_walk(const WalkArgs &a, int sign, int r, int u, uint32_t grid, uint32_t block,
                       hipStream_t st);
hipError_t launch_filter(const FilterArgs &a, uint32_t grid, hipStream_t st);
hipError_t launch_sieve(const SieveArgs &a, hipStream_t st);
// Sieve and compact one launch of a round (RoundSieveArgs): the primes go to
// out (ascending per piece, pieces in order) from *count_d on; at most cap
// entries of out are written, *count_d counts them all.  For each piece i of
// the launch, off[i] receives the pool offset of its first prime and psum[i]
// the sum of its primes mod 2^64.  blockcount and blocksum: one entry per
// kScanBlock words of the bitmap.
hipError_t launch_round_sieve(const RoundSieveArgs &a, uint32_t *blockcount, unsigned long long *blocksum,
                              uint64_t *out, uint32_t cap, uint32_t *count_d, uint32_t *off,
                              unsigned long long *psum, hipStream_t st);
// Compaction: primes of the chunk to out (ascending), appended at *count_d;
// at most cap entries of out are written, *count_d counts them all; *sum_d
// accumulates their sum mod 2^64 through blocksum (one entry per kScanBlock
// words of the chunk).
hipError_t launch_compact64(const uint32_t *bitmap, uint64_t lo, uint64_t nbits, uint32_t *blockcount,
                            unsigned long long *blocksum, uint64_t *out, uint32_t cap, uint32_t *count_d,
                            unsigned long long *sum_d, hipStream_t st);
hipError_t launch_compact32(const uint32_t *bitmap, uint64_t lo, uint64_t nbits, uint32_t *blockcount,
                            uint32_t *out, uint32_t cap, uint32_t *count_d, hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);
hipError_t launch_zero_u64(unsigned long long *p, uint32_t n, hipStream_t st);
hipError_t launch_state(const uint64_t *P, uint32_t n, uint32_t e, uint32_t nwin, uint32_t s, uint64_t *out,
                        hipStream_t st);

}  // namespace ppsd
