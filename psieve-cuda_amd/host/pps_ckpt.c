/* pps_ckpt.c -- the checkpoint record reader (see pps_ckpt.h). */
#include "pps_ckpt.h"

#include <string.h>

/* One unsigned 64-bit value: digits of base 10, or of base 16 after "0x";
 * at least one digit, no sign, no overflow.  Returns the end, or NULL. */
static const char *scan_value(const char *s, uint64_t *out) {
  unsigned base = 10;
  uint64_t v = 0;
  int digits = 0;
  if (s[0] == '0' && (s[1] == 'x' || s[1] == 'X')) {
    base = 16;
    s += 2;
  }
  for (;; s++, digits++) {
    unsigned d;
    const char ch = *s;
    if (ch >= '0' && ch <= '9')
      d = (unsigned)(ch - '0');
    else if (base == 16 && ch >= 'a' && ch <= 'f')
      d = (unsigned)(ch - 'a' + 10);
    else if (base == 16 && ch >= 'A' && ch <= 'F')
      d = (unsigned)(ch - 'A' + 10);
    else
      break;
    if (v > (UINT64_MAX - d) / base) return NULL;
    v = v * base + d;
  }
  if (!digits) return NULL;
  *out = v;
  return s;
}

int ckpt_parse_record(const char *line, const char *const names[], int n, uint64_t vals[]) {
  const char *s = line;
  for (int i = 0; i < n; i++) {
    const size_t len = strlen(names[i]);
    if (i > 0 && *s++ != ',') return 0;
    if (strncmp(s, names[i], len) != 0 || s[len] != '=') return 0;
    s = scan_value(s + len + 1, &vals[i]);
    if (!s) return 0;
  }
  if (*s == '\r') s++;
  if (*s == '\n') s++;
  return *s == '\0';
}

int ckpt_read_record(FILE *f, const char *const names[], int n, uint64_t vals[]) {
  char line[512];
  if (!fgets(line, sizeof line, f)) return 0;
  if (!strchr(line, '\n') && !feof(f)) return 0; /* longer than any record */
  return ckpt_parse_record(line, names, n, vals);
}
