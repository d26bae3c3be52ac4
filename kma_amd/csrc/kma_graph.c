/*
 * kma_graph.c - host side of the graph-building application benchmark (the
 * reference's tb_clTree workload: tb_clTree.c reads `{source,sink}` edges
 * and runs clTree_test, clTree_test_cache and clTree_test_al on KMA and on
 * the bump allocator).  No dataset ships with the reference, so
 * kma_graph_generate makes one; kma_graph_build uploads an edge list, times
 * one build launch and checks the graph against the edge list on the device.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_containers.h"
#include "kma_internal.h"
#include "kma_launch.h"
#include "kma_work.h"

/* Node ids: a bijection of 0..2^32-1 (odd multiplier, then xor), so the
 * nnodes ids are distinct and spread over the key space (a random insertion
 * order keeps the insert-only tree shallow). */
static uint32_t node_id(uint32_t i)
{
	return (i * 0x9E3779B1u) ^ 0x5BD1E995u;
}

static uint64_t splitmix(uint64_t *s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

int kma_graph_generate(uint32_t nnodes, uint64_t nedges, uint64_t seed, uint32_t *edges)
{
	if (nnodes == 0 || (nedges && !edges)) {
		set_err("bad arguments");
		return -1;
	}
	uint64_t s = seed;
	for (uint64_t e = 0; e < nedges; e++) {
		edges[2 * e] = node_id((uint32_t)(splitmix(&s) % nnodes));
		edges[2 * e + 1] = node_id((uint32_t)(splitmix(&s) % nnodes));
	}
	return 0;
}

static int cmp_u32(const void *a, const void *b)
{
	uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return x < y ? -1 : x > y;
}

static uint64_t find_key(const uint32_t *keys, uint64_t nk, uint32_t k)
{
	uint64_t lo = 0, hi = nk;
	while (lo < hi) {
		uint64_t mid = (lo + hi) / 2;
		if (keys[mid] < k)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

#define DEV_ALLOC(ptr, bytes)                                                        \
	do {                                                                         \
		if (hipMalloc((void **)&(ptr), (bytes) > 0 ? (bytes) : 1) != hipSuccess) { \
			set_err("device allocation of %llu bytes failed",            \
				(unsigned long long)(bytes));                        \
			goto done;                                                   \
		}                                                                    \
	} while (0)

int kma_graph_build(kma_heap *heap, kma_pma *pma, const uint32_t *edges, uint64_t n, int mode, uint32_t nwg,
		    int teardown, kma_graph_stats *out)
{
	if (!out || (n && !edges) || mode < KMA_GRAPH_PLAIN || mode > KMA_GRAPH_COLLECTIVE ||
	    (!pma && (!heap || heap->cfg.layout != KMA_LAYOUT_NATIVE)) ||
	    (pma && (mode != KMA_GRAPH_CACHE || teardown))) {
		set_err("bad arguments (native heap for every form; the bump allocator for the cache form only, "
			"without teardown)");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	if (kma_device_ok() != 0)
		return -1;
	if (nwg == 0) {
		uint64_t w = (n + 255) / 256;
		nwg = w < 1 ? 1 : (w > 1024 ? 1024 : (uint32_t)w);
	}
	/* expected graph: distinct keys, out-degree and sum of sink ids per key */
	uint32_t *keys = NULL, *deg = NULL;
	uint64_t *ssum = NULL, nk = 0;
	uint32_t *d_edges = NULL, *d_keys = NULL, *d_deg = NULL;
	uint64_t *d_ssum = NULL, *d_tree = NULL, *d_nodes = NULL;
	unsigned long long *d_stats = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	int rc = -1;
	keys = (uint32_t *)malloc((2 * n + 1) * 4);
	if (!keys) {
		set_err("host allocation failed");
		goto done;
	}
	if (n)
		memcpy(keys, edges, 2 * n * 4);
	qsort(keys, 2 * n, 4, cmp_u32);
	for (uint64_t i = 0; i < 2 * n; i++)
		if (i == 0 || keys[i] != keys[nk - 1])
			keys[nk++] = keys[i];
	deg = (uint32_t *)calloc(nk + 1, 4);
	ssum = (uint64_t *)calloc(nk + 1, 8);
	if (!deg || !ssum) {
		set_err("host allocation failed");
		goto done;
	}
	for (uint64_t e = 0; e < n; e++) {
		uint64_t k = find_key(keys, nk, edges[2 * e]);
		deg[k]++;
		ssum[k] += edges[2 * e + 1];
	}
	DEV_ALLOC(d_edges, 2 * n * 4);
	DEV_ALLOC(d_keys, nk * 4);
	DEV_ALLOC(d_deg, nk * 4);
	DEV_ALLOC(d_ssum, nk * 8);
	DEV_ALLOC(d_nodes, nk * 8);
	DEV_ALLOC(d_tree, 8);
	DEV_ALLOC(d_stats, 16 * 8);
	if ((n && hipMemcpy(d_edges, edges, 2 * n * 4, hipMemcpyHostToDevice) != hipSuccess) ||
	    (nk && (hipMemcpy(d_keys, keys, nk * 4, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(d_deg, deg, nk * 4, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(d_ssum, ssum, nk * 8, hipMemcpyHostToDevice) != hipSuccess)) ||
	    hipMemset(d_tree, 0, 8) != hipSuccess || hipMemset(d_stats, 0, 16 * 8) != hipSuccess ||
	    hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		set_err("graph setup failed");
		goto done;
	}
	void *dheap = heap ? heap->dev : NULL;
	void *dpma = pma ? kma_pma_device(pma) : NULL;
	hipEventRecord(e0, NULL);
	if (n && kmad_graph_build(dheap, dpma, d_tree, d_edges, n, mode, nwg, d_stats, NULL) != 0) {
		set_err("graph build launch failed");
		goto done;
	}
	hipEventRecord(e1, NULL);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("graph build kernel failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	if (kmad_graph_check(d_tree, d_keys, d_deg, d_ssum, nk, d_nodes, d_stats + 8, NULL) != 0 ||
	    (teardown && kmad_graph_free(dheap, d_nodes, d_deg, nk, NULL) != 0) || hipDeviceSynchronize() != hipSuccess) {
		set_err("graph check kernels failed");
		goto done;
	}
	unsigned long long st[16];
	if (hipMemcpy(st, d_stats, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("download failed");
		goto done;
	}
	out->gpu_ms = ms;
	out->edges = st[0];
	out->nodes = nk;
	out->node_mallocs = st[1];
	out->node_frees = st[2];
	out->link_mallocs = st[3];
	out->nulls = st[4];
	out->bad = st[5];
	out->violations = st[8];
	out->links_found = st[9];
	out->nwg = nwg;
	rc = 0;
done:
	free(keys);
	free(deg);
	free(ssum);
	if (d_edges)
		hipFree(d_edges);
	if (d_keys)
		hipFree(d_keys);
	if (d_deg)
		hipFree(d_deg);
	if (d_ssum)
		hipFree(d_ssum);
	if (d_nodes)
		hipFree(d_nodes);
	if (d_tree)
		hipFree(d_tree);
	if (d_stats)
		hipFree(d_stats);
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	return rc;
}
