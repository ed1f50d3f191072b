// ppsd_batch.h -- device data of the batched (many work units) walk.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppsd_kernels.h"

namespace ppsd {

// One piece (p-subrange) of a work unit, as the persistent kernel sees it.
struct PieceDesc {
  uint64_t kmin, kmax;
  uint32_t nmin, nmax;  // nmax exclusive
  uint32_t s, nwin;     // window and windows over [nmin, nmax)
  int32_t sign;         // +1, -1, 0 both
  uint32_t kstep, koffset;
  uint32_t prime_off, nprimes;  // its primes in the pool
  uint32_t unit;                // caller's unit index
  int32_t pre;                  // pre-test form (ppsd_walk.h)
};

// A unit of the work queue: count (<= 64*R) primes from pool offset
// pieces[piece].prime_off + first, windows w0 .. w0+nw of that piece
// (nw = 0: nothing to do).
struct Task {
  uint32_t piece, first, count, w0, nw;
};

// A class of equal tasks of one piece: nf blocks of `count` primes at
// first = f0 + i*fstep, times nseg segments of nw windows at w0 + j*nw;
// task t_begin + i*nseg + j of the round's order is block i, segment j.
struct TaskClass {
  uint64_t t_begin;
  uint32_t piece, f0, fstep, nf, count, w0, nw, nseg;
};

// The queue of one launch, built on the device from the round's classes
// (ascending t_begin): tasks [t0, t1) of the round's order, except that the
// last `tail` of them are each cut into up to kTailParts segments of at
// least tail_min_windows windows and queued after the rest, one part of
// every tail task in turn.  Writes the tasks and their number.
struct ExpandArgs {
  const TaskClass *classes;
  uint32_t nclasses;
  uint64_t t0, t1;
  uint32_t tail, parts, tail_min_windows;
  Task *out;
  uint32_t *nout;
};

// Raw hit of the batched walk.
struct BHit {
  uint64_t p;
  uint64_t k;
  uint32_t n;
  int32_t c;
  uint32_t piece;
  uint32_t pad;
};

struct BatchArgs {
  const uint64_t *pool;
  const PieceDesc *pieces;
  const Task *tasks;
  const uint32_t *ntasks;  // device: tasks in this round
  uint32_t *next_task;     // device: queue head, zeroed before each round
  BHit *hits;
  unsigned long long *hit_count;
  uint64_t hit_cap;
  BHit *kept;
  unsigned long long *kept_count;
  uint64_t kept_cap;
  const SmallPrime *small_primes;  // padded to a multiple of 64
  const uint16_t *pow2;            // their powers of two (ppsd_filter.h kPowPer each)
  uint32_t nsmall;
};

hipError_t launch_persist(const BatchArgs &a, int r, int u, uint32_t grid, hipStream_t st);
hipError_t launch_batch_filter(const BatchArgs &a, uint32_t grid, hipStream_t st);
hipError_t launch_emit(const BHit *src, uint64_t n, const PieceDesc *pieces, BHit *dst, hipStream_t st);
// Tasks of the launch described by a (ExpandArgs): (t1 - t0 - tail) + tail * parts of them.
hipError_t launch_expand(const ExpandArgs &a, hipStream_t st);

}  // namespace ppsd
