/* pps_hostsieve.h -- host segmented sieve of Eratosthenes over odd numbers,
 * the prime supply of `ppsieve --host-primes` (the role of the reference's
 * sieve.c:168-608 and main.c:736-832 when primes travel to the device). */
#ifndef PPS_HOSTSIEVE_H
#define PPS_HOSTSIEVE_H

#include <stddef.h>
#include <stdint.h>

typedef struct pps_hostsieve pps_hostsieve;

/* q[0..nq): the odd primes 3, 5, 7, ... ascending, at least up to
 * isqrt(p1 - 1) of every range later sieved.  The sieve keeps a pointer to q.
 * A range of at least nthreads * 2^26 numbers is sieved by nthreads threads
 * (the caller's and nthreads - 1 helpers). */
pps_hostsieve *pps_hostsieve_new(const uint32_t *q, uint32_t nq, int nthreads);
void pps_hostsieve_free(pps_hostsieve *h);

/* Every prime p in [p0, p1) (p1 <= 2^62), ascending, handed to emit() on
 * the calling thread in runs of at most 2^21.  Returns 0, or -1 if memory
 * ran out. */
int pps_hostsieve_range(pps_hostsieve *h, uint64_t p0, uint64_t p1,
                        void (*emit)(void *arg, const uint64_t *primes, size_t n), void *arg);

#endif
