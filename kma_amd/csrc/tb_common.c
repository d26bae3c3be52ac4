/*
 * tb_common.c - options, thread-configuration files and timing output of
 * the bench programs.  Option letters, their arguments, the file format and
 * the error messages follow the reference's test/cl.c:190-307 (options
 * match on their first two characters, as there); -c has no counterpart
 * here (the allocator runs on the GPU only) and is refused; -g prints the
 * device counters of each run instead of compiling kernels with debug
 * information (they are built ahead of time); -l is new: it lists the runs
 * without touching the GPU.
 */
#include "tb_common.h"

#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static tb_tuple one_item = {1, 1, 1};
tb_options tb_opt = {0, 0, 0, 1, &one_item};

/* Thread-configuration files come from users; cap what one may ask for. */
#define TB_MAX_ENTRIES 65536u

int tb_threadcfg_read(const char *file)
{
	FILE *fp = fopen(file, "r");
	if (!fp) {
		printf("Error: file %s not found\n\n", file);
		return -1;
	}
	int n = 0;
	if (fscanf(fp, "%d\n", &n) < 1 || n <= 0 || (unsigned)n > TB_MAX_ENTRIES) {
		printf("Error: malformed thread config file provided\n\n");
		fclose(fp);
		return -1;
	}
	tb_tuple *t = (tb_tuple *)calloc((size_t)n, sizeof(*t));
	if (!t) {
		fclose(fp);
		return -1;
	}
	for (int i = 0; i < n; i++) {
		/* %SCNu64 skips blanks before each number, so "{8, 8, 4}" reads */
		if (fscanf(fp, "{%" SCNu64 ",%" SCNu64 ",%" SCNu64 "}\n", &t[i].x, &t[i].y, &t[i].z) != 3) {
			printf("Error: malformed thread config file provided\n\n");
			free(t);
			fclose(fp);
			return -1;
		}
	}
	fclose(fp);
	tb_opt.wi_entries = (unsigned)n;
	tb_opt.wi = t;
	return 0;
}

int tb_options_read(int argc, char **argv, tb_extra_fn extra)
{
	for (int i = 1; i < argc; i++) {
		const char *a = argv[i];
		if (strncmp(a, "-c", 2) == 0) {
			printf("Error: -c: this build has no CPU device; the allocator runs on the GPU only\n\n");
			return -1;
		} else if (strncmp(a, "-p", 2) == 0) {
			tb_opt.print_buffer = 1;
		} else if (strncmp(a, "-g", 2) == 0) {
			tb_opt.debug = 1;
		} else if (strncmp(a, "-l", 2) == 0) {
			tb_opt.list = 1;
		} else if (strncmp(a, "-t", 2) == 0) {
			if (++i >= argc) {
				printf("Error: -t requires a {x,y,z} tuple\n\n");
				return -1;
			}
			tb_tuple t;
			if (sscanf(argv[i], "%" SCNu64 ",%" SCNu64 ",%" SCNu64, &t.x, &t.y, &t.z) != 3)
				return -1;
			one_item = t;
			tb_opt.wi_entries = 1;
			tb_opt.wi = &one_item;
		} else if (strncmp(a, "-T", 2) == 0) {
			if (++i >= argc) {
				printf("Error: -T requires a file name\n\n");
				return -1;
			}
			if (tb_threadcfg_read(argv[i]) < 0)
				return -1;
		} else {
			int used = extra ? extra(i, argc, argv) : -1;
			if (used < 0) {
				printf("Error: Unknown option: %s\n\n", a);
				return -1;
			}
			i += used;
		}
	}
	return 0;
}

void tb_options_print(void)
{
	printf("  -c           run on the host CPU (not available: this build has no CPU device)\n");
	printf("  -p           hex-dump the first superblocks of each heap after its runs\n");
	printf("  -g           after each run, print the device counters and any give-ups\n");
	printf("  -l           only list the planned runs; the GPU is not opened\n");
	printf("  -t x,y,z     run one launch shape: x*y*z work-items\n");
	printf("  -T file      run every launch shape listed in `file`\n");
	printf("\nA shape file starts with the number of shapes, followed by one {x,y,z}\n");
	printf("line per shape.  Without -t or -T a single work-item runs.\n");
}

uint64_t tb_threads(const tb_tuple *t)
{
	const uint64_t cap = 1ull << 40;
	if (t->x == 0 || t->y == 0 || t->z == 0)
		return 0;
	if (t->x > cap || t->y > cap / t->x || t->z > cap / (t->x * t->y))
		return UINT64_MAX;
	return t->x * t->y * t->z;
}

void tb_print_times(const uint64_t *us, unsigned n)
{
	uint64_t lo = n ? us[0] : 0, hi = lo, tot = 0;
	for (unsigned i = 0; i < n; i++) {
		lo = us[i] < lo ? us[i] : lo;
		hi = us[i] > hi ? us[i] : hi;
		tot += us[i];
	}
	uint64_t avg = n ? tot / n : 0;
	printf("Avg: %" PRIu64 ".%06" PRIu64 " ( %" PRIu64 ".%06" PRIu64 " - %" PRIu64 ".%06" PRIu64 " ) Total: %" PRIu64
	       ".%06" PRIu64 "\n",
	       avg / 1000000, avg % 1000000, lo / 1000000, lo % 1000000, hi / 1000000, hi % 1000000, tot / 1000000,
	       tot % 1000000);
}

void tb_dump_words(const void *buf, uint64_t bytes, uint64_t group)
{
	const uint8_t *b = (const uint8_t *)buf;
	for (uint64_t off = 0; off + 4 <= bytes; off += 4) {
		if (group && off % group == 0)
			printf("\n");
		if (off % 16 == 0)
			printf("\n%08" PRIx64 ": ", off);
		uint32_t w;
		memcpy(&w, b + off, 4);
		printf("%08x ", w);
	}
	printf("\n");
}
