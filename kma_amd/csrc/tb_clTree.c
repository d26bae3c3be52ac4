/*
 * tb_clTree.c - the `clTree` bench program: the reference's graph-building
 * benchmark (test/tb_clTree.c:59-489) on libkma.so.  It reads a data set
 * (the number of entries, then one `{source,sink}` line per edge;
 * tb_clTree.c:350-389) or generates one (-G), and for every thread
 * configuration builds the graph on a fresh heap in each of the
 * reference's forms, in its order:
 *     KMA-1:          clTree_test        one kma_malloc per lane
 *                     clTree_test_cache  a lost node kept for the lane's next insert
 *                     clTree_test_al     work-group collective allocation
 *     Poormans heap:  clTree_test_cache  on the bump allocator
 * printing "Threads: <n>" and the timing line per configuration.  Each
 * build is checked on the device against the edge list (every id in the
 * tree once, with its out-degree and sum of sink ids); a failed check is
 * reported and makes the exit status non-zero.  Lanes are launched in
 * work-groups of 256: a configuration of x*y*z work-items runs on
 * ceil(x*y*z / 256) work-groups (the reference lets the runtime pick the
 * work-group size), the edges striding over all lanes.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_containers.h"
#include "kma_work.h"
#include "tb_common.h"

#define MAX_EDGES (1u << 28)
#define MAX_WG 65536u

static const char *dataset = NULL;
static uint32_t gen_nodes = 0, gen_seed = 1;
static uint64_t gen_edges = 0;
static uint64_t sblocks = 0; /* -H: native heap superblocks (0: sized from the data set) */
static uint32_t nctx = 4096; /* -C: native heap allocation contexts */
static unsigned samples = 1; /* -s */

static int tree_opts(int i, int argc, char **argv)
{
	const char *a = argv[i];
	if (a[0] != '-') {
		dataset = a;
		return 0;
	}
	if (i + 1 >= argc)
		return -1;
	if (strncmp(a, "-G", 2) == 0) {
		unsigned long long n = 0, e = 0;
		unsigned s = 1;
		int got = sscanf(argv[i + 1], "%llu,%llu,%u", &n, &e, &s);
		if (got < 2 || n == 0 || n > UINT32_MAX || e == 0 || e > MAX_EDGES)
			return -1;
		gen_nodes = (uint32_t)n;
		gen_edges = e;
		gen_seed = got == 3 ? s : 1u;
		return 1;
	}
	char *end;
	unsigned long long v = strtoull(argv[i + 1], &end, 0);
	if (*argv[i + 1] == 0 || *argv[i + 1] == '-' || *end != 0)
		return -1;
	if (strncmp(a, "-H", 2) == 0 && v >= 1 && v <= (1u << 24)) {
		sblocks = v;
		return 1;
	}
	if (strncmp(a, "-C", 2) == 0 && v >= 8 && v <= 65536 && (v & (v - 1)) == 0) {
		nctx = (uint32_t)v;
		return 1;
	}
	if (strncmp(a, "-s", 2) == 0 && v >= 1 && v <= 1000) {
		samples = (unsigned)v;
		return 1;
	}
	return -1;
}

static void usage(void)
{
	printf("Usage: clTree [options] [edge file]\n");
	printf("Builds a graph's adjacency lists in an insert-only tree with each\n");
	printf("allocation form, one build per launch shape.\n");
	tb_options_print();
	printf("  -G n,e[,s]   generate e edges over n node ids (seed s) instead of a file\n");
	printf("  -H n         superblocks of 16 KiB per heap (default: sized from the edges)\n");
	printf("  -C n         allocation contexts per heap, a power of two (default 4096)\n");
	printf("  -s n         timed builds per launch shape (default 1)\n");
	printf("\nAn edge file holds the edge count on its first line, then one\n");
	printf("{source,sink} pair per line.\n");
}

/* The data set file (tb_clTree.c:350-389); *edges = 2 * n u32 on success. */
static int read_dataset(uint32_t **edges, uint64_t *n)
{
	if (!dataset) {
		printf("Error: No data set file specified\n");
		return -1;
	}
	FILE *fp = fopen(dataset, "r");
	if (!fp) {
		printf("Error: specified file cannot be opened (%s)\n", dataset);
		return -1;
	}
	unsigned cnt = 0;
	if (fscanf(fp, "%u\n", &cnt) < 1 || cnt > MAX_EDGES) {
		printf("Error: Dataset file malformed\n");
		fclose(fp);
		return -1;
	}
	uint32_t *e = (uint32_t *)malloc(2ull * (cnt ? cnt : 1) * 4);
	if (!e) {
		fclose(fp);
		return -1;
	}
	for (unsigned i = 0; i < cnt; i++)
		if (fscanf(fp, "{%u,%u}\n", &e[2 * i], &e[2 * i + 1]) < 2) {
			printf("Error: Dataset file malformed\n");
			free(e);
			fclose(fp);
			return -1;
		}
	fclose(fp);
	*edges = e;
	*n = cnt;
	return 0;
}

static const struct {
	const char *section, *name;
	int mode, pma;
} forms[] = {
	{"KMA-1:", "clTree_test", KMA_GRAPH_PLAIN, 0},
	{NULL, "clTree_test_cache", KMA_GRAPH_CACHE, 0},
	{NULL, "clTree_test_al", KMA_GRAPH_COLLECTIVE, 0},
	{"Poormans heap:", "clTree_test_cache", KMA_GRAPH_CACHE, 1},
};

static int run_form(unsigned f, const uint32_t *edges, uint64_t n)
{
	int rc = 0;
	if (forms[f].section)
		printf("%s\n", forms[f].section);
	printf("-- Executing %s --\n", forms[f].name);
	uint64_t *us = (uint64_t *)calloc(samples, sizeof(uint64_t));
	/* nodes <= 2 edges: at most 64 B per node (48-B nodes in the 64-B
	 * class) and 16 B per edge live, plus one cached node per lane; the
	 * heap adds a superblock per context for each of the two classes
	 * (nodes, links) */
	uint64_t need = 144ull * n;
	for (unsigned t = 0; t < tb_opt.wi_entries && us; t++) {
		uint64_t threads = tb_threads(&tb_opt.wi[t]);
		if (threads == 0 || threads > (uint64_t)MAX_WG * 256u) {
			printf("Error: %" PRIu64 " threads: outside 1..%u\n", threads, MAX_WG * 256u);
			rc = -1;
			continue;
		}
		uint32_t nwg = (uint32_t)((threads + 255) / 256);
		uint64_t bytes = need + 64ull * 256u * nwg;
		uint64_t nsb = sblocks ? sblocks : 1024u + 2u * (bytes >> 14) + 2u * nctx;
		if (tb_opt.list) {
			printf("Threads: %-5" PRIu64 " (%u work-groups of 256)\n", threads, nwg);
			continue;
		}
		kma_graph_stats st;
		memset(&st, 0, sizeof(st));
		for (unsigned s = 0; s < samples; s++) {
			kma_heap *h = NULL;
			kma_pma *p = NULL;
			if (forms[f].pma) {
				p = kma_pma_create(2 * bytes + (2u << 20), 16);
			} else {
				kma_config c;
				memset(&c, 0, sizeof(c));
				c.layout = KMA_LAYOUT_NATIVE;
				c.sb_log2 = 14;
				c.sblocks = nsb;
				c.nctx = nctx;
				h = kma_create_ex(&c);
			}
			int ok = (h || p) && kma_graph_build(h, p, edges, n, forms[f].mode, nwg, 0, &st) == 0;
			if (!ok)
				printf("Error: Could not execute kernel: %s\n", kma_last_error());
			else if (st.violations || st.bad || st.nulls || st.edges != n || st.links_found != n) {
				printf("Error: graph check failed: %" PRIu64 " violations, %" PRIu64 " of %" PRIu64
				       " links found, %" PRIu64 " NULLs (heap exhausted: raise -H)\n",
				       st.violations, st.links_found, n, st.nulls);
				ok = 0;
			}
			if (h)
				kma_destroy(h);
			if (p)
				kma_pma_destroy(p);
			if (!ok) {
				rc = -1;
				break;
			}
			us[s] = (uint64_t)(st.gpu_ms * 1000.0 + 0.5);
		}
		if (rc)
			break;
		printf("Threads: %-5" PRIu64 " ", threads);
		tb_print_times(us, samples);
		if (tb_opt.debug)
			printf("\tedges %" PRIu64 " nodes %" PRIu64 " node_mallocs %" PRIu64 " node_frees %" PRIu64
			       " link_mallocs %" PRIu64 " nulls %" PRIu64 " links_found %" PRIu64 " edges/s %.4g\n",
			       st.edges, st.nodes, st.node_mallocs, st.node_frees, st.link_mallocs, st.nulls,
			       st.links_found, st.gpu_ms > 0 ? (double)st.edges / (st.gpu_ms * 1e-3) : 0.0);
	}
	free(us);
	printf("\n");
	return rc;
}

int main(int argc, char **argv)
{
	/* one line at a time, so a run that ends abnormally keeps its lines */
	setvbuf(stdout, NULL, _IOLBF, 0);
	printf("Tree test using the gfx950 device allocator\n");
	if (tb_options_read(argc, argv, tree_opts)) {
		usage();
		return -1;
	}
	uint32_t *edges = NULL;
	uint64_t n = 0;
	if (gen_edges) {
		edges = (uint32_t *)malloc(2 * gen_edges * 4);
		if (!edges || kma_graph_generate(gen_nodes, gen_edges, gen_seed, edges) != 0) {
			printf("Error: could not generate the data set\n");
			return 1;
		}
		n = gen_edges;
	} else if (read_dataset(&edges, &n) < 0) {
		return -1;
	}
	if (!tb_opt.list && kma_device_ok() != 0) {
		printf("Error: no usable GPU: %s\n", kma_last_error());
		free(edges);
		return 1;
	}
	int rc = 0;
	for (unsigned f = 0; f < sizeof(forms) / sizeof(forms[0]); f++)
		rc |= run_form(f, edges, n);
	free(edges);
	return rc ? 1 : 0;
}
