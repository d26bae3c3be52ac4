/*
 * kma_graph.hip - gfx950 kernels of the graph-building application benchmark
 * (the reference's tb_clTree workload, tb_clTree.cl:12-273).  Every edge
 * {source, sink} makes sure both of its nodes are in an insert-only tree
 * keyed by node id (kma_tree.h), allocates a link and pushes it onto the
 * source node's adjacency list.  Nodes and links come from
 *   - one kma_malloc per lane, a node that lost its insert race being freed
 *     (tb_clTree.cl:12-70, clTree_test);
 *   - the same with the lost node kept for the lane's next insert
 *     (tb_clTree.cl:87-173, clTree_test_cache), on the native heap or on the
 *     bump allocator (the reference's "poormans heap" run);
 *   - work-group collective kma_wg_malloc calls: one for the source nodes
 *     missing from the tree, one for the sink nodes, one for the links
 *     (tb_clTree.cl:175-273, clTree_test_al, which slices ArrayList pages).
 * Every node carries its adjacency list as an embedded linked queue
 * (kma_lq, kma_lqueue.h: the reference's per-node clqueue that every link is
 * enqueued to, tb_clTree.cl:67, 168, 271); a link's first word is its queue
 * link.
 */
#include <hip/hip_runtime.h>

#include "kma_device.h"
#include "kma_launch.h"
#include "kma_lqueue.h"
#include "kma_pma.h"
#include "kma_tree.h"
#include "kma_wg_device.h"

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

namespace {

constexpr uint32_t GBLOCK = 256;

/* A graph node is a tree node followed by the queue of its links. */
struct gnode {
	kma_tree_node t;
	kma_lq adj; /* of glink */
};
struct glink {
	uint64_t next; /* queue link (kma_lqueue.h) */
	uint64_t sink; /* gnode * */
};
static_assert(sizeof(gnode) == 48 && sizeof(glink) == 16, "graph record sizes");

/* counters at stats[k], summed per wavefront */
enum : uint32_t { G_EDGES, G_NODE_MALLOC, G_NODE_FREE, G_LINK_MALLOC, G_NULLS, G_BAD, G_N };

/* Named fields, not an array: the compiler merges increments on different
 * paths into one with a computed index, which would put an array in scratch. */
struct gcount {
	uint32_t edges = 0, node_malloc = 0, node_free = 0, link_malloc = 0, nulls = 0, bad = 0;
};

__device__ __forceinline__ void flush(unsigned long long *stats, const gcount &c)
{
	const uint32_t v[G_N] = {c.edges, c.node_malloc, c.node_free, c.link_malloc, c.nulls, c.bad};
#pragma unroll
	for (uint32_t k = 0; k < G_N; k++) {
		uint64_t x = v[k];
		for (int o = 32; o > 0; o >>= 1)
			x += __shfl_xor(x, o);
		if (__lane_id() == 0 && x)
			atomicAdd(stats + k, (unsigned long long)x);
	}
}

__device__ __forceinline__ void st64(uint64_t *p, uint64_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Enqueue link l (to sink) on node src's adjacency queue; false if the
 * queue refused it (a corrupt queue). */
__device__ __forceinline__ bool adj_push(gnode *src, glink *l, gnode *sink)
{
	st64(&l->sink, (uint64_t)(uintptr_t)sink);
	return kma_lq_push(&src->adj, l) != 0;
}

/* The links of a node no call is using, in queue order: from the node at
 * the head, the queue's stub skipped; at most `limit` of them. */
template <class F>
__device__ __forceinline__ uint32_t adj_walk(gnode *nd, uint32_t limit, F &&f)
{
	uint32_t d = 0;
	const uint64_t stub = (uint64_t)(uintptr_t)&nd->adj.stub;
	for (uint64_t w = kma_lq_ptr(nd->adj.head); w && d <= limit;) {
		const uint64_t nx = kma_lq_ptr(*(const uint64_t *)(uintptr_t)w);
		if (w != stub) {
			f((glink *)(uintptr_t)w);
			d++;
		}
		w = nx;
	}
	return d;
}

/* The node with `key`: found in the tree, or `cache` inserted.  On a lost
 * insert race the winner's node is found on the next look-up and `cache`
 * stays with the caller.  NULL if a node was needed and cache was NULL
 * (the caller allocates and calls again). */
__device__ __forceinline__ gnode *ensure_from(kma_tree *t, uint32_t key, gnode *&cache, uint32_t *lost)
{
	for (uint32_t attempt = 0; attempt < 4; attempt++) {
		gnode *n = (gnode *)kma_tree_get(t, key);
		if (n)
			return n;
		if (!cache)
			return nullptr;
		kma_lq_init(&cache->adj);
		cache->t.key = key;
		if (kma_tree_add(t, &cache->t)) {
			n = cache;
			cache = nullptr;
			return n;
		}
	}
	*lost = 1; /* a lost race whose winner is not found */
	return nullptr;
}

/* Per-lane allocation from the native heap (H) or the bump allocator (P). */
struct lane_alloc {
	kma_dheap *H;
	kma_pma_dheap *P;
	__device__ __forceinline__ void *get(size_t size) { return P ? pma_malloc(P, size) : kma_malloc(H, size); }
	__device__ __forceinline__ void put(void *p)
	{
		if (!P)
			kma_free(H, p);
	}
};

/* The node with `key` for a per-lane form, allocating (and, in the plain
 * form, freeing a node that lost its insert race) as needed.  The counts are
 * returned as values for the caller to add: increments of different fields
 * on different paths were merged by the compiler into one increment through
 * a selected address, which put the counters in scratch. */
struct ensured {
	gnode *n;
	uint32_t mallocs, nulls, frees, bad;
};

__device__ __forceinline__ ensured ensure_lane(kma_tree *t, uint32_t key, gnode *&cache, bool keep, lane_alloc &a)
{
	uint32_t mallocs = 0, frees = 0;
	for (uint32_t attempt = 0; attempt < 4; attempt++) {
		gnode *n = (gnode *)kma_tree_get(t, key);
		if (n)
			return ensured{n, mallocs, 0u, frees, 0u};
		if (!cache) {
			cache = (gnode *)a.get(sizeof(gnode));
			mallocs++;
			if (!cache)
				return ensured{nullptr, mallocs, 1u, frees, 0u};
		}
		kma_lq_init(&cache->adj);
		cache->t.key = key;
		if (kma_tree_add(t, &cache->t)) {
			n = cache;
			cache = nullptr;
			return ensured{n, mallocs, 0u, frees, 0u};
		}
		if (!keep) {
			/* plain form: the node that lost the insert race is freed */
			a.put(cache);
			frees++;
			cache = nullptr;
		}
	}
	return ensured{nullptr, mallocs, 0u, frees, 1u}; /* lost races whose winners were not found */
}

} // namespace

/* Per-lane forms: grid-stride over the edges (tb_clTree.cl:57-67). */
__global__ void __launch_bounds__(GBLOCK) k_graph_lane(kma_dheap *H, kma_pma_dheap *P, kma_tree *t,
						       const uint32_t *edges, uint64_t n, uint32_t keep,
						       unsigned long long *stats)
{
	lane_alloc a = {H, P};
	gcount c;
	gnode *cache = nullptr;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		ensured e = ensure_lane(t, edges[2 * i], cache, keep != 0, a);
		c.node_malloc += e.mallocs;
		c.nulls += e.nulls;
		c.node_free += e.frees;
		c.bad += e.bad;
		gnode *src = e.n;
		if (!src)
			break;
		e = ensure_lane(t, edges[2 * i + 1], cache, keep != 0, a);
		c.node_malloc += e.mallocs;
		c.nulls += e.nulls;
		c.node_free += e.frees;
		c.bad += e.bad;
		gnode *snk = e.n;
		if (!snk)
			break;
		glink *l = (glink *)a.get(sizeof(glink));
		c.link_malloc++;
		c.nulls += l ? 0u : 1u;
		if (!l)
			break;
		if (adj_push(src, l, snk))
			c.edges++;
		else
			c.bad++;
	}
	if (cache) {
		a.put(cache);
		c.node_free += P ? 0u : 1u;
	}
	flush(stats, c);
}

/* Work-group collective form (tb_clTree.cl:175-273): the loop runs the same
 * number of iterations in every thread, and each iteration makes three
 * collective allocations (source nodes, sink nodes, links) for the lanes
 * that need one. */
__global__ void __launch_bounds__(GBLOCK) k_graph_wg(kma_dheap *H, kma_tree *t, const uint32_t *edges, uint64_t n,
						     unsigned long long *stats)
{
	__shared__ kma_wg_lds<GBLOCK> L;
	kma_wg_lds_init(L);
	gcount c;
	gnode *cache = nullptr;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	const uint64_t span = (n + stride - 1) / stride * stride;
	bool dead = false; /* this lane met a NULL: it takes part in the collectives only */
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < span; i += stride) {
		const bool has = i < n && !dead;
		const uint32_t ks = has ? edges[2 * i] : 0u, kk = has ? edges[2 * i + 1] : 0u;
		gnode *src = has ? (gnode *)kma_tree_get(t, ks) : nullptr;
		bool need = has && !src && !cache;
		void *p = kma_wg_malloc<GBLOCK, 16u>(H, need ? sizeof(gnode) : 0u, L);
		if (need)
			cache = (gnode *)p;
		c.node_malloc += need ? 1u : 0u;
		c.nulls += need && !p ? 1u : 0u;
		uint32_t lost = 0;
		if (has && !src)
			src = ensure_from(t, ks, cache, &lost);
		gnode *snk = has && src ? (gnode *)kma_tree_get(t, kk) : nullptr;
		need = has && src && !snk && !cache;
		p = kma_wg_malloc<GBLOCK, 16u>(H, need ? sizeof(gnode) : 0u, L);
		if (need)
			cache = (gnode *)p;
		c.node_malloc += need ? 1u : 0u;
		c.nulls += need && !p ? 1u : 0u;
		if (has && src && !snk)
			snk = ensure_from(t, kk, cache, &lost);
		c.bad += lost;
		const bool linkit = has && src && snk;
		glink *l = (glink *)kma_wg_malloc<GBLOCK, 16u>(H, linkit ? sizeof(glink) : 0u, L);
		const bool pushed = linkit && l && adj_push(src, l, snk);
		c.link_malloc += linkit ? 1u : 0u;
		c.edges += pushed ? 1u : 0u;
		c.bad += linkit && l && !pushed ? 1u : 0u;
		c.nulls += linkit && !l ? 1u : 0u;
		dead = dead || (i < n && !(linkit && l));
	}
	c.node_free += cache ? 1u : 0u;
	kma_wg_free<GBLOCK, 16u>(H, cache, L);
	flush(stats, c);
}

/* Check against the edge list: key i (sorted, distinct) must be in the tree
 * with exactly deg[i] links whose sinks' keys sum to ssum[i].  At most
 * deg[i] + 1 links are walked, so a corrupt list cannot loop.  nodes[i]
 * receives the node (for the teardown). */
__global__ void __launch_bounds__(GBLOCK) k_graph_check(kma_tree *t, const uint32_t *keys, const uint32_t *deg,
							const uint64_t *ssum, uint64_t nk, uint64_t *nodes,
							unsigned long long *stats)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t bad = 0, links = 0;
	if (i < nk) {
		gnode *nd = (gnode *)kma_tree_get(t, keys[i]);
		nodes[i] = (uint64_t)(uintptr_t)nd;
		if (!nd) {
			bad = 1;
		} else {
			/* the reference's queue validator (tb_clQueue.c:126-157) on the
			 * node's queue: the chain from head ends, at the node tail
			 * names, passing the stub once; then the links' sinks */
			uint64_t s = 0, last = 0, w = kma_lq_ptr(nd->adj.head);
			const uint64_t stub = (uint64_t)(uintptr_t)&nd->adj.stub;
			uint32_t d = 0, stubs = 0;
			while (w && d <= deg[i] && stubs <= 1u) {
				if (w == stub) {
					stubs++;
				} else {
					s += ((const gnode *)(uintptr_t)((const glink *)(uintptr_t)w)->sink)->t.key;
					d++;
				}
				last = w;
				w = kma_lq_ptr(*(const uint64_t *)(uintptr_t)w);
			}
			links = d;
			bad = w != 0 || stubs != 1u || last != kma_lq_ptr(nd->adj.tail) || d != deg[i] || s != ssum[i];
		}
	}
	uint64_t x = bad, y = links;
	for (int o = 32; o > 0; o >>= 1) {
		x += __shfl_xor(x, o);
		y += __shfl_xor(y, o);
	}
	if (__lane_id() == 0) {
		if (x)
			atomicAdd(stats + 0, (unsigned long long)x);
		if (y)
			atomicAdd(stats + 1, (unsigned long long)y);
	}
}

/* Teardown on the native heap: free node i's links and node i (the nodes
 * were recorded by the check, so no tree walk runs during the frees). */
__global__ void __launch_bounds__(GBLOCK) k_graph_free(kma_dheap *H, const uint64_t *nodes, const uint32_t *deg,
						       uint64_t nk)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nk || !nodes[i])
		return;
	gnode *nd = (gnode *)(uintptr_t)nodes[i];
	adj_walk(nd, deg[i], [&](glink *lk) { kma_free(H, lk); });
	kma_free(H, nd);
}

extern "C" int kmad_graph_build(void *heap, void *pma, void *tree, const uint32_t *d_edges, uint64_t n, int mode,
				uint32_t nwg, unsigned long long *d_stats, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	if (mode == 2)
		hipLaunchKernelGGL(k_graph_wg, dim3(nwg), dim3(GBLOCK), 0, s, (kma_dheap *)heap, (kma_tree *)tree,
				   d_edges, n, d_stats);
	else
		hipLaunchKernelGGL(k_graph_lane, dim3(nwg), dim3(GBLOCK), 0, s, (kma_dheap *)heap,
				   (kma_pma_dheap *)pma, (kma_tree *)tree, d_edges, n, (uint32_t)(mode == 1), d_stats);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_graph_check(void *tree, const uint32_t *d_keys, const uint32_t *d_deg, const uint64_t *d_ssum,
				uint64_t nk, uint64_t *d_nodes, unsigned long long *d_stats, void *stream)
{
	if (nk == 0)
		return 0;
	hipLaunchKernelGGL(k_graph_check, dim3((uint32_t)((nk + GBLOCK - 1) / GBLOCK)), dim3(GBLOCK), 0,
			   (hipStream_t)stream, (kma_tree *)tree, d_keys, d_deg, d_ssum, nk, d_nodes, d_stats);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_graph_free(void *heap, const uint64_t *d_nodes, const uint32_t *d_deg, uint64_t nk, void *stream)
{
	if (nk == 0)
		return 0;
	hipLaunchKernelGGL(k_graph_free, dim3((uint32_t)((nk + GBLOCK - 1) / GBLOCK)), dim3(GBLOCK), 0,
			   (hipStream_t)stream, (kma_dheap *)heap, d_nodes, d_deg, nk);
	KMA_CHECK(hipGetLastError());
	return 0;
}
