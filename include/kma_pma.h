/*
 * kma_pma.h - device side of the bump ("poor man's") allocator, HIP for
 * gfx950.  Header-only.
 *
 * Same semantics as the reference's PMA (pma.cl:32-48): malloc takes the
 * next `size` bytes of the region with one atomic add on the head and
 * returns NULL when the request does not fit; free is a no-op; a failed
 * request still advances the head (it is never rolled back).  Differences,
 * none of which changes a result the reference defines:
 *   - the head is 64-bit, so failed requests cannot wrap it (the
 *     reference's u32 head, pma.h:56, wraps after 4 GiB of requests,
 *     pma.cl:37);
 *   - `bytes` is the data capacity, the header is outside it (the
 *     reference counts the header in `bytes`, pma.c:37, and checks the
 *     offset against it, pma.cl:38, so it can overrun by the header's size);
 *   - sizes are rounded up to the heap's alignment (1 = the reference's
 *     unaligned behaviour);
 *   - the lanes of a wavefront that call together are served by ONE atomic
 *     add of their summed sizes (the head is the single hottest word of this
 *     allocator); each lane gets the offset it would have got had the lanes'
 *     adds run one after another in lane order.
 */
#ifndef KMA_PMA_H
#define KMA_PMA_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KMA_PMA_MAGIC 0x31414d504b4d414bull /* "KAMKPMA1" */
#define KMA_PMA_HDR 256u                    /* data starts here */

typedef struct kma_pma_dheap {
	uint64_t magic;
	uint64_t bytes; /* data capacity */
	uint64_t align; /* power of two >= 1 */
	uint64_t pad0[13];
	uint64_t head; /* next free data offset (own 128-B line) */
} kma_pma_dheap;

__device__ inline void *pma_malloc(kma_pma_dheap *H, size_t size)
{
	uint64_t a = H->align;
	uint64_t sz = (size + a - 1) & ~(a - 1);
	uint64_t act = __ballot(1);
	uint32_t me = __lane_id();
	uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
	/* exclusive prefix of the sizes over the calling lanes, in lane order */
	uint64_t pre = 0, tot = 0;
	for (uint64_t m = act; m; m &= m - 1ull) {
		uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
		uint64_t v = __shfl(sz, (int)l);
		if (l == me)
			pre = tot;
		tot += v;
	}
	uint64_t base = 0;
	if (me == leader)
		base = __hip_atomic_fetch_add(&H->head, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	base = __shfl(base, (int)leader);
	uint64_t off = base + pre;
	if (size == 0 || off + sz > H->bytes)
		return nullptr;
	return (char *)H + KMA_PMA_HDR + off;
}

/* pma.cl:44-48: free is a no-op. */
__device__ inline void pma_free(kma_pma_dheap *H, void *p)
{
	(void)H;
	(void)p;
}

#endif /* KMA_PMA_H */
