/* pps_input.c -- sieve-file input (-i) for the ppsieve host: the ABCD and
 * NewPGen formats of pps/app.c:188-517.
 *
 * NewPGen: a header "<p>:P:<x>:2:<t>" (t = 1 or 257 for k*2^n+1, -1 or 255
 * for k*2^n-1; pps/app.c:222-224), then "<k> <n>" pairs.
 * ABCD: blocks, each a header "ABCD <k>*2^$a<+|->1 [<n>]" followed by lines
 * holding the increments of n for that k (pps/app.c:242-302, 443-505).
 * The twin-prime build reads "<p>:T:<x>:2:3" NewPGen files and ABCD blocks
 * "ABCD (6*$a+3)*2^<n>+1 & (6*$a+3)*2^<n>-1 [<a>]" whose lines hold
 * increments of a (k = 6a+3) at a fixed n (pps/app.c:209-212, 260-275).
 *
 * The header formats and the meaning of each term are the reference's; the
 * reader is written line by line and skips blank lines and terms outside
 * the search ranges (where the reference would write outside its bitmap).
 */
#define _GNU_SOURCE
#include "pps_input.h"

#include <ctype.h>
#include <inttypes.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(char *err, size_t errlen, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (err && errlen) vsnprintf(err, errlen, fmt, ap);
  va_end(ap);
  return -1;
}

static int abcd_header(const char *line, uint64_t *k, int *sign, unsigned int *n) {
  char ch = 0;
  if (sscanf(line, "ABCD %" SCNu64 "*2^$a%c1 [%u]", k, &ch, n) != 3) return 0;
  *sign = ch == '+' ? 1 : -1;
  return 1;
}

static int newpgen_header(const char *line, int *sign) {
  uint64_t p0;
  int t;
  if (sscanf(line, "%" SCNu64 ":P:%*c:2:%d", &p0, &t) != 2) return 0;
  if (t != 1 && t != -1 && t != 255 && t != 257) return 0;
  *sign = (int)(signed char)t; /* 255 -> -1, 257 -> +1 */
  return 1;
}

/* Twin-prime forms: k = 6a+3 (the header's a), both signs, one n per block. */
static int abcd_twin_header(const char *line, uint64_t *a, unsigned int *n) {
  unsigned int m;
  return sscanf(line, "ABCD (6*$a+3)*2^%u+1 & (6*$a+3)*2^%u-1 [%" SCNu64 "]", &m, n, a) == 3 && m == *n;
}

static int newpgen_twin_header(const char *line) {
  uint64_t p0;
  char ch;
  return sscanf(line, "%" SCNu64 ":T:%*c:2:%c", &p0, &ch) == 2 && ch == '3';
}

static int blank(const char *s) {
  while (*s && isspace((unsigned char)*s)) s++;
  return *s == 0;
}

/* Walk every term (k, n) of the file after its header line, calling visit.
 * visit returns nonzero to stop with an error (message already in err). */
typedef int (*visit_fn)(void *arg, uint64_t k, unsigned int n, unsigned int line);

static int walk_abcd(FILE *f, char **buf, size_t *cap, uint64_t k, unsigned int n, visit_fn visit, void *arg,
                     unsigned int *line) {
  int sign;
  if (visit(arg, k, n, *line)) return -1;
  while (getline(buf, cap, f) > 0) {
    ++*line;
    const char *s = *buf;
    if (isdigit((unsigned char)s[0])) {
      n += (unsigned int)strtoul(s, NULL, 10);
      if (visit(arg, k, n, *line)) return -1;
    } else if (abcd_header(s, &k, &sign, &n)) {
      if (visit(arg, k, n, *line)) return -1;
    } else if (!blank(s)) {
      break; /* the reference stops reading at the first line that is neither */
    }
  }
  return 0;
}

/* Twin ABCD: per block a fixed n and increments of a (pps/app.c:260-275). */
static int walk_abcd_twin(FILE *f, char **buf, size_t *cap, uint64_t a, unsigned int n, visit_fn visit, void *arg,
                          unsigned int *line) {
  if (visit(arg, 6 * a + 3, n, *line)) return -1;
  while (getline(buf, cap, f) > 0) {
    ++*line;
    const char *s = *buf;
    if (isdigit((unsigned char)s[0])) {
      a += strtoull(s, NULL, 10);
      if (visit(arg, 6 * a + 3, n, *line)) return -1;
    } else if (abcd_twin_header(s, &a, &n)) {
      if (visit(arg, 6 * a + 3, n, *line)) return -1;
    } else if (!blank(s)) {
      break;
    }
  }
  return 0;
}

static int walk_newpgen(FILE *f, visit_fn visit, void *arg) {
  uint64_t k;
  unsigned int n, line = 0; /* pairs counted from 1, as pps/app.c:386-391 */
  while (fscanf(f, " %" SCNu64 " %u", &k, &n) == 2) {
    ++line;
    if (visit(arg, k, n, line)) return -1;
  }
  return 0;
}

struct scan_state {
  pps_input_scan *s;
  int any;
};

static int scan_visit(void *arg, uint64_t k, unsigned int n, unsigned int line) {
  struct scan_state *st = arg;
  pps_input_scan *s = st->s;
  (void)line;
  if (!st->any) {
    s->k0 = s->k1 = k;
    s->n0 = n;
    s->n1 = n;
    st->any = 1;
  }
  if (k < s->k0) s->k0 = k;
  if (k > s->k1) s->k1 = k;
  if (n < s->n0) s->n0 = n;
  if (n > s->n1) s->n1 = n;
  return 0;
}

int pps_input_scan_file(const char *fn, int twin, pps_input_scan *out, char *err, size_t errlen) {
  FILE *f = fopen(fn, "r");
  char *buf = NULL;
  size_t cap = 0;
  unsigned int line = 1;
  uint64_t k;
  unsigned int n;
  int rc = 0;
  struct scan_state st = {out, 0};
  if (!f) return fail(err, errlen, "%s: cannot open", fn);
  memset(out, 0, sizeof *out);
  if (getline(&buf, &cap, f) <= 0) {
    rc = fail(err, errlen, "Invalid header in input file `%s'", fn);
  } else if (twin && abcd_twin_header(buf, &k, &n)) {
    out->format = PPS_INPUT_ABCD;
    out->sign = 1;
    rc = walk_abcd_twin(f, &buf, &cap, k, n, scan_visit, &st, &line);
  } else if (twin && newpgen_twin_header(buf)) {
    out->format = PPS_INPUT_NEWPGEN;
    out->sign = 1;
    rc = walk_newpgen(f, scan_visit, &st);
    if (!rc && !st.any) rc = fail(err, errlen, "Invalid line 2 in input file `%s'", fn);
  } else if (twin) {
    rc = fail(err, errlen, "Invalid header in input file `%s'", fn);
  } else if (abcd_header(buf, &k, &out->sign, &n)) {
    out->format = PPS_INPUT_ABCD;
    rc = walk_abcd(f, &buf, &cap, k, n, scan_visit, &st, &line);
  } else if (newpgen_header(buf, &out->sign)) {
    out->format = PPS_INPUT_NEWPGEN;
    rc = walk_newpgen(f, scan_visit, &st);
    if (!rc && !st.any) rc = fail(err, errlen, "Invalid line 2 in input file `%s'", fn);
  } else {
    rc = fail(err, errlen, "Invalid header in input file `%s'", fn);
  }
  if (!rc && ferror(f)) rc = fail(err, errlen, "Error reading input file `%s'", fn);
  out->n1 += 1; /* one past the largest n, as pps/app.c:317 */
  free(buf);
  fclose(f);
  return rc;
}

struct read_state {
  uint64_t kmin, kmax;
  unsigned int nmin, nmax, kstep, koffset;
  int check_residue; /* NewPGen: k must be koffset mod kstep */
  int count_all;     /* ABCD: every term read counts (pps/app.c:459), NewPGen: terms set */
  uint8_t *bitmap;
  uint64_t row_bytes;
  unsigned int count;
  char *err;
  size_t errlen;
  const char *fn;
};

static int read_visit(void *arg, uint64_t k, unsigned int n, unsigned int line) {
  struct read_state *st = arg;
  if (st->check_residue && k % st->kstep != st->koffset)
    return fail(st->err, st->errlen, "Invalid line %u in input file `%s'", line, st->fn);
  if (st->count_all) st->count++;
  if (k < st->kmin || k > st->kmax || n < st->nmin || n > st->nmax) return 0;
  const uint64_t bit = (k - st->kmin) / st->kstep;
  if (bit / 8 >= st->row_bytes) return 0;
  st->bitmap[(uint64_t)(n - st->nmin) * st->row_bytes + bit / 8] |= (uint8_t)(1u << (bit % 8));
  if (!st->count_all) st->count++;
  return 0;
}

int pps_input_read_file(const char *fn, int format, int twin, uint64_t kmin, uint64_t kmax, unsigned int nmin,
                        unsigned int nmax, unsigned int kstep, unsigned int koffset, uint8_t *bitmap,
                        uint64_t row_bytes, unsigned int *count, char *err, size_t errlen) {
  FILE *f = fopen(fn, "r");
  char *buf = NULL;
  size_t cap = 0;
  unsigned int line = 1;
  uint64_t k;
  unsigned int n;
  int sign, rc = 0;
  struct read_state st = {kmin, kmax, nmin, nmax, kstep, koffset, format == PPS_INPUT_NEWPGEN,
                          format == PPS_INPUT_ABCD, bitmap, row_bytes, 0, err, errlen, fn};
  if (!f) return fail(err, errlen, "%s: cannot open", fn);
  if (getline(&buf, &cap, f) <= 0) {
    rc = fail(err, errlen, "Invalid header in input file `%s'", fn);
  } else if (twin && format == PPS_INPUT_ABCD && abcd_twin_header(buf, &k, &n)) {
    rc = walk_abcd_twin(f, &buf, &cap, k, n, read_visit, &st, &line);
  } else if (twin && format == PPS_INPUT_NEWPGEN && newpgen_twin_header(buf)) {
    rc = walk_newpgen(f, read_visit, &st);
  } else if (!twin && format == PPS_INPUT_ABCD && abcd_header(buf, &k, &sign, &n)) {
    rc = walk_abcd(f, &buf, &cap, k, n, read_visit, &st, &line);
  } else if (!twin && format == PPS_INPUT_NEWPGEN && newpgen_header(buf, &sign)) {
    rc = walk_newpgen(f, read_visit, &st);
  } else {
    rc = fail(err, errlen, "Invalid header in input file `%s'", fn);
  }
  if (!rc && ferror(f)) rc = fail(err, errlen, "Error reading input file `%s'", fn);
  if (count) *count = st.count;
  free(buf);
  fclose(f);
  return rc;
}
