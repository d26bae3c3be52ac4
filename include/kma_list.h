/*
 * kma_list.h - device side of the growable array list, HIP for gfx950.
 * Header-only.
 *
 * Same structure and call surface as the reference's clArrayList
 * (clArrayList.cl:32-216, clArrayList.h:44-62): the list is a queue of pages
 * taken from a heap; a page is a 24-byte header {next, count, taken}
 * followed by `count` objects of obj_size bytes.
 *   kma_list_grow(l, objs)        one page of objs objects, its data pointer
 *                                 (NULL if the heap cannot serve it);
 *   kma_list_grow_local(l, objs)  work-group collective: every thread asks
 *                                 for objs (maybe 0) objects, the work-group
 *                                 takes ONE page for the sum and each thread
 *                                 gets data + (exclusive prefix of objs) *
 *                                 obj_size, NULL if it asked for none or the
 *                                 page could not be had;
 *   kma_list_get(l, i)            object i, counting through the pages in
 *                                 queue order;
 *   kma_list_clear(l)             dequeue every page and free it.
 * The reference's work-group scan is a Blelloch up/down sweep over a
 * 128-entry local array with 2 log2(WG) barriers (clArrayList.cl:111-176);
 * here each wavefront scans in registers and one LDS pass combines the
 * wavefront totals (two barriers, any work-group size up to BLOCK).
 * Pages come from the native heap (kma_malloc) and go through the kma_queue
 * ring; `next` is kept in the page header for layout parity and is 0.
 */
#ifndef KMA_LIST_H
#define KMA_LIST_H

#include "kma_device.h"
#include "kma_queue.h"

#define KMA_LIST_MAGIC 0x314c414d4b4d414bull /* "KAMKMAL1" */
#define KMA_LIST_PAGE_HDR 24u

typedef struct kma_dlist {
	uint64_t magic;
	uint64_t obj_size;
	kma_dheap *heap;   /* device heap the pages come from */
	kma_dqueue *queue; /* the pages, in grow order */
} kma_dlist;

typedef struct kma_list_page {
	uint64_t next; /* unused by the ring queue; 0 */
	uint64_t count;
	uint64_t taken;
} kma_list_page;

template <int BLOCK>
struct kma_list_lds {
	uint64_t wsum[BLOCK / 64];
	uint64_t base;
};

__device__ inline void *kma_list_grow(kma_dlist *l, uint32_t objs)
{
	uint64_t size = (uint64_t)objs * l->obj_size + KMA_LIST_PAGE_HDR;
	kma_list_page *p = (kma_list_page *)kma_malloc(l->heap, size);
	if (!p)
		return nullptr;
	p->next = 0;
	p->count = objs;
	p->taken = objs;
	if (!kma_enqueue(l->queue, p)) {
		kma_free(l->heap, p);
		return nullptr;
	}
	return (char *)p + KMA_LIST_PAGE_HDR;
}

/* Work-group collective (every thread of the 1-D work-group calls it). */
template <int BLOCK>
__device__ inline void *kma_list_grow_local(kma_dlist *l, uint32_t objs, kma_list_lds<BLOCK> &L)
{
	const uint32_t tid = threadIdx.x, nth = blockDim.x, w = tid >> 6, nw = (nth + 63u) >> 6;
	const uint32_t me = __lane_id();
	/* inclusive scan within the wavefront (inactive tail lanes add 0) */
	uint64_t v = objs, x = v;
	uint64_t act = __ballot(1);
	for (uint32_t o = 1; o < 64; o <<= 1) {
		uint64_t y = __shfl_up(x, o);
		if (me >= o && ((act >> (me - o)) & 1ull))
			x += y;
	}
	uint32_t last = 63u - (uint32_t)__clzll((long long)act);
	if (me == last)
		L.wsum[w] = x;
	__syncthreads();
	if (tid == 0) {
		uint64_t acc = 0;
		for (uint32_t k = 0; k < nw; k++) {
			uint64_t t = L.wsum[k];
			L.wsum[k] = acc;
			acc += t;
		}
		void *d = acc ? kma_list_grow(l, (uint32_t)acc) : nullptr;
		L.base = (uint64_t)(uintptr_t)d;
	}
	__syncthreads();
	uint64_t pre = L.wsum[w] + x - v;
	uint64_t b = L.base;
	__syncthreads();
	if (objs == 0 || b == 0)
		return nullptr;
	return (char *)(uintptr_t)b + pre * l->obj_size;
}

/* Object i (clArrayList.cl:178-200); NULL past the end.  Not safe against
 * concurrent grow or clear. */
__device__ inline void *kma_list_get(kma_dlist *l, uint64_t i)
{
	using namespace kma_q;
	kma_dqueue *q = l->queue;
	uint64_t h = ld(&q->head), t = ld(&q->tail), cml = 0;
	for (uint64_t k = h; k < t; k++) {
		kma_list_page *p = (kma_list_page *)(uintptr_t)ld(&slot(q, k)->ptr);
		uint64_t c = p->count;
		if (cml + c > i)
			return (char *)p + KMA_LIST_PAGE_HDR + (i - cml) * l->obj_size;
		cml += c;
	}
	return nullptr;
}

/* Dequeue every page and free it (clArrayList.cl:206-216); safe to call from
 * many threads at once while nobody uses the list. */
__device__ inline void kma_list_clear(kma_dlist *l)
{
	void *p = kma_dequeue(l->queue);
	while (p) {
		kma_free(l->heap, p);
		p = kma_dequeue(l->queue);
	}
}

#endif /* KMA_LIST_H */
