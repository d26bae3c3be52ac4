/*
 * tb_common.h - command-line options, thread-configuration files and the
 * timing line shared by the bench programs (kma, clTree), with the
 * reference's option letters, file format and messages (test/cl.c:190-307,
 * test/timing.c:107-123).
 */
#ifndef TB_COMMON_H
#define TB_COMMON_H

#include <stdint.h>

typedef struct tb_tuple {
	uint64_t x, y, z;
} tb_tuple;

typedef struct tb_options {
	int print_buffer;  /* -p */
	int debug;         /* -g: per-run device counters */
	int list;          /* -l: print the planned runs, touch no GPU */
	unsigned wi_entries;
	tb_tuple *wi;      /* thread configurations (-t or -T; default {1,1,1}) */
} tb_options;

extern tb_options tb_opt;

/* A program's own options: called with argv[i] not one of the common
 * ones; returns how many following arguments it consumed, or -1. */
typedef int (*tb_extra_fn)(int i, int argc, char **argv);

/* Parse argv; 0, or -1 after printing the error (the caller prints usage). */
int tb_options_read(int argc, char **argv, tb_extra_fn extra);

/* Read a thread-configuration file: a count, then that many {x,y,z} lines. */
int tb_threadcfg_read(const char *file);

/* The common options' usage text. */
void tb_options_print(void);

/* x*y*z, or UINT64_MAX when that exceeds 2^40. */
uint64_t tb_threads(const tb_tuple *t);

/* "Avg: s.us ( min - max ) Total: s.us" of n samples given in microseconds. */
void tb_print_times(const uint64_t *us, unsigned n);

/* Hex dump of 32-bit words, four to a line with the byte offset, a blank
 * line before each `group` bytes. */
void tb_dump_words(const void *buf, uint64_t bytes, uint64_t group);

#endif
