/* pps_ckpt.h -- reading the text records of the checkpoint file.
 *
 * A checkpoint file is two lines of comma-separated name=value fields, the
 * framework's ("pmin=..,p=..,count=..,sum=0x..,checksum=0x..") and the
 * application's ("nmin=..,nmax=..,factor_count=.."); the formats are those
 * the reference writes (main.c:305-330, pps/app.c:1348-1352), so that
 * checkpoints move between the two programs.
 */
#ifndef PPS_CKPT_H
#define PPS_CKPT_H

#include <stdint.h>
#include <stdio.h>

/* Parses one record: exactly the n fields names[0..n) in that order, each
 * "name=value" with a decimal value or a hexadecimal one after "0x", fields
 * separated by single commas, optionally ended by a newline.  On success
 * vals[0..n) receive the values and 1 is returned; otherwise 0, with vals
 * unspecified. */
int ckpt_parse_record(const char *line, const char *const names[], int n, uint64_t vals[]);

/* Reads the next line of f and parses it as ckpt_parse_record does; 0 when
 * there is no line, or it is too long for a record, or it does not parse. */
int ckpt_read_record(FILE *f, const char *const names[], int n, uint64_t vals[]);

#endif
