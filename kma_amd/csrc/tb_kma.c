/*
 * tb_kma.c - the `kma` bench program: the reference's allocator benchmark
 * (test/tb_kma.c:222-384) on libkma.so.  For each of the reference's three
 * test kernels (kma_test_malloc, _lowvar, _highvar; kma.cl:424-526) it
 * creates a heap (kma_create(127) as tb_kma.c:234 does, or with -N a
 * native heap: by default 1 GiB of 16 KiB superblocks with 4,096 allocation
 * contexts, one per wavefront resident at once), then for every thread
 * configuration and iteration count runs ONE launch of x*y*z work-items,
 * each doing `iters` malloc/write/free
 * rounds (the kernel's loop, a NULL ends the work-item), and prints one
 * line per run in the reference's format:
 *     <threads> threads <iters> iters: Avg: s.us ( min - max ) Total: s.us
 * Times are the launch's device time (HIP events on its stream); the
 * reference takes host wall-clock time around enqueue + finish.  The
 * work-item id is the flattened 3-D global id, as in the kernels, so each
 * work-item requests the same sizes as there.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_work.h"
#include "tb_common.h"

static unsigned iters = 100, step = 0; /* -i [step,]iters (tb_kma.c:322-347) */
static int native = 0;                 /* -N: native layout */
static uint64_t sblocks = 0;           /* -H: superblocks (0: 127, or 65536 native) */
static uint32_t nctx = 0;              /* -C: native allocation contexts (0: 4096) */
static unsigned block = 256;           /* -b: work-group size */
static unsigned samples = 1;           /* -s: timed launches per point */
static unsigned deadline_ms = 5000;    /* -D: reference-layout deadline per launch */
static const char *only = NULL;        /* -k: one kernel */

#define TB_MAX_THREADS (1ull << 26)

static const struct {
	const char *name;
	uint32_t mode;
} kernels[] = {
	{"kma_test_malloc", KMA_WL_REF_TEST},
	{"kma_test_malloc_lowvar", KMA_WL_LOWVAR},
	{"kma_test_malloc_highvar", KMA_WL_HIGHVAR},
};

static int parse_uint(const char *s, uint64_t lo, uint64_t hi, uint64_t *out)
{
	char *end;
	unsigned long long v = strtoull(s, &end, 0);
	if (*s == 0 || *end != 0 || *s == '-' || v < lo || v > hi)
		return -1;
	*out = v;
	return 0;
}

static int kma_opts(int i, int argc, char **argv)
{
	const char *a = argv[i];
	uint64_t v;
	if (strncmp(a, "-i", 2) == 0) {
		unsigned s, it;
		if (i + 1 >= argc)
			return -1;
		switch (sscanf(argv[i + 1], "%u,%u", &s, &it)) {
		case 1:
			iters = s;
			step = 0;
			break;
		case 2:
			iters = it;
			step = s;
			break;
		default:
			return -1;
		}
		printf("Step: %u\n", step);
		printf("Iter: %u\n", iters);
		return 1;
	}
	if (strncmp(a, "-N", 2) == 0) {
		native = 1;
		return 0;
	}
	if (i + 1 >= argc)
		return -1;
	if (strncmp(a, "-H", 2) == 0 && parse_uint(argv[i + 1], 1, 1u << 24, &sblocks) == 0)
		return 1;
	if (strncmp(a, "-C", 2) == 0 && parse_uint(argv[i + 1], 8, 65536, &v) == 0 && (v & (v - 1)) == 0) {
		nctx = (uint32_t)v;
		return 1;
	}
	if (strncmp(a, "-b", 2) == 0 && parse_uint(argv[i + 1], 64, 256, &v) == 0 && v % 64 == 0) {
		block = (unsigned)v;
		return 1;
	}
	if (strncmp(a, "-s", 2) == 0 && parse_uint(argv[i + 1], 1, 1000, &v) == 0) {
		samples = (unsigned)v;
		return 1;
	}
	if (strncmp(a, "-D", 2) == 0 && parse_uint(argv[i + 1], 0, 3600000, &v) == 0) {
		deadline_ms = (unsigned)v;
		return 1;
	}
	if (strncmp(a, "-k", 2) == 0) {
		for (unsigned k = 0; k < sizeof(kernels) / sizeof(kernels[0]); k++)
			if (strcmp(argv[i + 1], kernels[k].name) == 0) {
				only = kernels[k].name;
				return 1;
			}
	}
	return -1;
}

static void usage(void)
{
	printf("Usage: kma [options]\n");
	printf("Times the allocator's three test kernels, one launch per launch shape and\n");
	printf("loop count.\n");
	printf("  -i [step,]n  loop count per work-item (default 100); with a step, sweep the\n");
	printf("               counts 2, step, 2*step, ... up to n\n");
	tb_options_print();
	printf("  -N           use the native heap layout (default: the reference layout with\n");
	printf("               4 KiB superblocks)\n");
	printf("  -H n         superblocks per heap (default 127; native 65536 of 16 KiB)\n");
	printf("  -C n         native allocation contexts, a power of two (default 4096)\n");
	printf("  -b n         work-group size, a multiple of 64 up to 256 (default 256)\n");
	printf("  -s n         timed launches per point (default 1)\n");
	printf("  -D ms        reference layout: waits give up after ms per launch, the\n");
	printf("               mallocs return NULL and the run line counts them (default\n");
	printf("               5000; 0: only the 4 s per-call watchdogs)\n");
	printf("  -k name      one kernel only: kma_test_malloc, kma_test_malloc_lowvar or\n");
	printf("               kma_test_malloc_highvar (default: all three, in that order)\n");
}

/* Iteration counts of one thread configuration (tb_kma.c:250-296): iters
 * alone without a step; with one, 2 first, then step, 2*step, ... (for
 * steps of 1 and 2: 2, 2 + step, ...), all up to iters. */
static uint64_t next_iters(uint64_t its)
{
	if (its == 2 && step > 2)
		return step;
	return its + step;
}

static kma_heap *make_heap(void)
{
	if (!native)
		return kma_create(sblocks ? (unsigned)sblocks : 127u);
	kma_config c;
	memset(&c, 0, sizeof(c));
	c.layout = KMA_LAYOUT_NATIVE;
	c.sb_log2 = 14;
	c.sblocks = sblocks ? sblocks : 65536u;
	c.nctx = nctx ? nctx : 4096u;
	return kma_create_ex(&c);
}

/* One kernel over every configuration; 0, or -1 after a failed run. */
static int run_kernel(unsigned k)
{
	int rc = 0;
	kma_heap *h = NULL;
	if (!tb_opt.list) {
		h = make_heap();
		if (!h) {
			printf("KMA_test: could not create the heap: %s\n", kma_last_error());
			return -1;
		}
	}
	printf("-- Executing %s --\n", kernels[k].name);
	uint64_t *us = (uint64_t *)calloc(samples, sizeof(uint64_t));
	for (unsigned t = 0; t < tb_opt.wi_entries && rc == 0; t++) {
		const tb_tuple *w = &tb_opt.wi[t];
		uint64_t threads = tb_threads(w);
		for (uint64_t its = step ? 2u : iters; its <= iters; its = next_iters(its)) {
			kma_run_stats st;
			memset(&st, 0, sizeof(st));
			memset(us, 0, samples * sizeof(uint64_t));
			if (tb_opt.list) {
				printf("%-5" PRIu64 " threads %-5" PRIu64 " iters: {%" PRIu64 ",%" PRIu64 ",%" PRIu64 "}\n",
				       threads, its, w->x, w->y, w->z);
			} else {
				for (unsigned s = 0; s < samples && its > 0; s++) {
					kma_workload wl;
					memset(&wl, 0, sizeof(wl));
					wl.mode = kernels[k].mode;
					wl.form = KMA_FORM_FUSED;
					wl.nwi = threads;
					wl.block = block;
					wl.rounds = (uint32_t)its;
					wl.deadline_ms = native ? 0 : deadline_ms;
					int ok = threads > 0 && threads <= TB_MAX_THREADS;
					if (!ok || kma_run_workload(h, &wl, &st) != 0) {
						printf("KMA_test: Could not execute %s on %" PRIu64 " threads: %s\n",
						       kernels[k].name, threads, ok ? kma_last_error() : "outside 1..2^26 work-items");
						rc = -1;
						break;
					}
					us[s] = (uint64_t)(st.gpu_ms * 1000.0 + 0.5);
				}
				if (rc)
					break;
				printf("%-5" PRIu64 " threads %-5" PRIu64 " iters: ", threads, its);
				tb_print_times(us, samples);
				if (!native && (st.deadline_hits || st.waits[0] || st.waits[1] || st.waits[2] || st.waits[3]))
					printf("\tgave up: deadline %" PRIu64 " reserve %" PRIu64 " claim %" PRIu64 " dequeue %" PRIu64
					       " enqueue %" PRIu64 " (nulls %" PRIu64 " of %" PRIu64 " mallocs)\n",
					       st.deadline_hits, st.waits[0], st.waits[1], st.waits[2], st.waits[3], st.nulls,
					       st.mallocs);
				if (tb_opt.debug) {
					printf("\tops %" PRIu64 " mallocs %" PRIu64 " nulls %" PRIu64 " bad %" PRIu64
					       " double_frees %" PRIu64 " bad_frees %" PRIu64 " waits %" PRIu64 " %" PRIu64
					       " %" PRIu64 " %" PRIu64 " ops/s %.4g\n",
					       st.ops, st.mallocs, st.nulls, st.bad, st.double_frees, st.bad_frees,
					       st.waits[0], st.waits[1], st.waits[2], st.waits[3],
					       st.gpu_ms > 0 ? (double)st.ops / (st.gpu_ms * 1e-3) : 0.0);
					{
						kma_heap_report r;
						if (kma_heap_check(h, &r) == 0)
							printf("\theap walk: %" PRIu64 " problems, %" PRIu64 " orphans, %" PRIu64
							       " live blocks\n%s",
							       r.problems, r.orphans, r.live_blocks, r.first);
					}
				}
			}
			if (step == 0)
				break;
		}
	}
	free(us);
	if (h && rc == 0 && tb_opt.print_buffer) {
		kma_config c;
		kma_get_config(h, &c);
		uint64_t sb = 1ull << c.sb_log2, n = 16 * sb;
		if (n > kma_heap_bytes(h))
			n = kma_heap_bytes(h);
		void *img = malloc(n);
		if (img && kma_snapshot(h, 0, img, n) == 0)
			tb_dump_words(img, n, sb);
		free(img);
	}
	if (h)
		kma_destroy(h);
	printf("\n");
	return rc;
}

int main(int argc, char **argv)
{
	/* one line at a time, so a run that ends abnormally keeps its lines */
	setvbuf(stdout, NULL, _IOLBF, 0);
	printf("KMA device allocator bench (gfx950)\n");
	if (tb_options_read(argc, argv, kma_opts)) {
		usage();
		return -1;
	}
	if (!tb_opt.list && kma_device_ok() != 0) {
		printf("Error: no usable GPU: %s\n", kma_last_error());
		return 1;
	}
	int rc = 0;
	for (unsigned k = 0; k < sizeof(kernels) / sizeof(kernels[0]); k++)
		if (!only || strcmp(only, kernels[k].name) == 0)
			rc |= run_kernel(k);
	return rc ? 1 : 0;
}
