/* pps_input.h -- sieve-file input (-i): the ABCD and NewPGen formats the
 * reference reads in pps/app.c:188-517, turned into the per-n candidate
 * bitmap its test_factor consults (pps/app.c:181-184). */
#ifndef PPS_INPUT_H
#define PPS_INPUT_H

#include <stddef.h>
#include <stdint.h>

#define PPS_INPUT_NEWPGEN 1
#define PPS_INPUT_ABCD 2

/* 1 in the twin-prime build: the twin headers "ABCD (6*$a+3)*2^<n>+1 &
 * (6*$a+3)*2^<n>-1 [<a>]" (increments of a follow) and "<p>:T:<x>:2:3"
 * (pps/app.c:209-212, 260-275, 414-490). */
#ifdef PPS_TWIN
#define PPS_INPUT_TWIN 1
#else
#define PPS_INPUT_TWIN 0
#endif

/* What a first pass over the file finds (pps/app.c:188-347). */
typedef struct pps_input_scan {
  int format;          /* PPS_INPUT_ABCD or PPS_INPUT_NEWPGEN */
  int sign;            /* +1 or -1, from the header */
  uint64_t k0, k1;     /* smallest and largest k */
  unsigned int n0, n1; /* smallest n, largest n + 1 */
} pps_input_scan;

/* Detect the format from the first line and scan every term.  Returns 0,
 * or -1 with a message in err (invalid header, unreadable file). */
int pps_input_scan_file(const char *fn, int twin, pps_input_scan *out, char *err, size_t errlen);

/* Read the terms into bitmap: rows for n in [nmin, nmax], row_bytes bytes
 * each; term (k, n) sets bit (k - kmin) / kstep of row n - nmin, as
 * pps/app.c:388-397 (NewPGen) and 443-505 (ABCD) do.  Terms outside
 * [kmin, kmax] x [nmin, nmax] are skipped.  A NewPGen line whose k is not
 * koffset mod kstep is an error (pps/app.c:389-393).  *count receives the
 * terms set.  Returns 0 or -1 with a message in err. */
int pps_input_read_file(const char *fn, int format, int twin, uint64_t kmin, uint64_t kmax, unsigned int nmin,
                        unsigned int nmax, unsigned int kstep, unsigned int koffset, uint8_t *bitmap,
                        uint64_t row_bytes, unsigned int *count, char *err, size_t errlen);

#endif
