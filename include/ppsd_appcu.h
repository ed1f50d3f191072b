/* ppsd_appcu.h -- the reference's CUDA-layer entry points, exported by
 * libppsd.so so that an unmodified PPSieve host (main.c + pps/app.c) links
 * against it in place of pps/appcu.cu.  Names, signatures and calling
 * contract are those of pps/appcu.h; behaviour follows pps/appcu.cu.
 *
 * These functions read the application's globals (pmin, pmax, nmin, nmax,
 * kmin, kmax, search_proth, thread_subsections, num_threads) and report
 * errors through its bmsg/bexit, exactly as the reference's device layer
 * does.  Those symbols are weak references inside libppsd.so: they bind to
 * the host program that defines them, and calling these entry points from a
 * program that does not define them fails with an error message.
 */
#ifndef PPSD_APPCU_H
#define PPSD_APPCU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pps/app.h:88-92: per worker, the n-subsection boundaries the host re-walk
 * uses; slot 0 holds nmax, slot 8 nmin. */
struct n_subsection {
  unsigned int n_subsection_start[9];
  unsigned int *first_n_subsection;
};

/* pps/appcu.h:31, pps/appcu.cu:155-209: window clamps and checks, ld_r0 and
 * ld_bbits for the host re-walk. */
void cuda_init(void);

/* pps/appcu.h:27, pps/appcu.cu:213-404: opens GPU gpuno for worker th,
 * fills thread_subsections[th]; returns the batch size (0 = GPU unusable). */
unsigned int cuda_app_init(int gpuno, int th, unsigned int cthread_count);

/* pps/appcu.h:29, pps/appcu.cu:1205-1261: starts the walk of cthread_count
 * primes (P may be NULL to re-run the previous batch). */
void check_ns(const uint64_t *P, const unsigned int cthread_count, const int th);

/* pps/appcu.h:30, pps/appcu.cu:1263-1308: waits, then sets bit j of
 * factor_found[i] when P[i] has a divisor k*2^n+c with n in subsection j. */
void get_factors_found(unsigned char *factor_found, const unsigned int cthread_count, const uint64_t start_t,
                       int *check_ns_delay);

/* pps/appcu.h:32, pps/appcu.cu:101-122: frees the calling worker's GPU state. */
void cuda_finalize(void);

/* pps/appcu.h:35-37 */
extern unsigned int ld_nstep;
extern int ld_bbits;
extern uint64_t ld_r0;

#ifdef __cplusplus
}
#endif
#endif /* PPSD_APPCU_H */
