/*
 * kma_containers.hip - gfx950 kernels behind the container test runners
 * (include/kma_containers.h): bump allocator, pointer queue, array list and
 * binary search tree.  The data-structure code itself is in the device
 * headers (kma_pma.h, kma_queue.h, kma_list.h, kma_tree.h).
 */
#include <hip/hip_runtime.h>

#include "kma_device.h"
#include "kma_launch.h"
#include "kma_list.h"
#include "kma_pma.h"
#include "kma_lqueue.h"
#include "kma_queue.h"
#include "kma_tree.h"

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

static inline uint32_t nblocks(uint64_t n, uint32_t t)
{
	return (uint32_t)((n + t - 1) / t);
}

/* ---------------- bump allocator ---------------- */
__global__ void k_pma_init(kma_pma_dheap *H, uint64_t bytes, uint64_t align)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		H->magic = KMA_PMA_MAGIC;
		H->bytes = bytes;
		H->align = align;
		H->head = 0;
	}
}

__global__ void k_pma_serial(kma_pma_dheap *H, const uint32_t *sizes, uint64_t n, uint64_t *out)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	char *d0 = (char *)H + KMA_PMA_HDR;
	for (uint64_t i = 0; i < n; i++) {
		char *p = (char *)pma_malloc(H, sizes[i]);
		out[i] = p ? (uint64_t)(p - d0) : ~0ull;
	}
}

__global__ void k_pma_parallel(kma_pma_dheap *H, const uint32_t *sizes, uint64_t n, uint64_t *out)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	char *d0 = (char *)H + KMA_PMA_HDR;
	char *p = (char *)pma_malloc(H, sizes[i]);
	out[i] = p ? (uint64_t)(p - d0) : ~0ull;
}

extern "C" int kmad_pma_init(void *h, uint64_t bytes, uint64_t align)
{
	hipLaunchKernelGGL(k_pma_init, dim3(1), dim3(64), 0, 0, (kma_pma_dheap *)h, bytes, align);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_pma_run(void *h, const uint32_t *sizes, uint64_t n, int serial, uint64_t *out)
{
	if (serial)
		hipLaunchKernelGGL(k_pma_serial, dim3(1), dim3(64), 0, 0, (kma_pma_dheap *)h, sizes, n, out);
	else
		hipLaunchKernelGGL(k_pma_parallel, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_pma_dheap *)h, sizes, n,
				   out);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ---------------- pointer queue ---------------- */
__global__ void k_queue_init(kma_dqueue *q, uint64_t cap)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) {
		q->magic = KMA_QUEUE_MAGIC;
		q->cap = cap;
		q->head = 0;
		q->tail = 0;
	}
	if (i < cap) {
		kma_dqueue_slot *s = (kma_dqueue_slot *)((char *)q + KMA_QUEUE_HDR) + i;
		s->seq = i;
		s->ptr = 0;
	}
}

__global__ void k_queue_enqueue(kma_dqueue *q, char *items, uint64_t n)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		kma_enqueue(q, items + 16u * i);
}

__global__ void k_queue_shuffle(kma_dqueue *q, char *items, uint64_t n, uint32_t rounds)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	kma_enqueue(q, items + 16u * i);
	for (uint32_t r = 0; r < rounds; r++) {
		void *it = kma_dequeue(q);
		if (it)
			kma_enqueue(q, it);
	}
}

__global__ void k_queue_serial(kma_dqueue *q, char *items, uint64_t n, int64_t *order)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	for (uint64_t i = 0; i < n; i++)
		kma_enqueue(q, items + 16u * i);
	uint64_t k = 0;
	for (;;) {
		char *it = (char *)kma_dequeue(q);
		if (!it)
			break;
		if (k < n)
			order[k] = (int64_t)((it - items) / 16);
		k++;
	}
	for (; k < n; k++)
		order[k] = -1;
}

extern "C" int kmad_queue_init(void *q, uint64_t cap)
{
	hipLaunchKernelGGL(k_queue_init, dim3(nblocks(cap, 256)), dim3(256), 0, 0, (kma_dqueue *)q, cap);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_queue_run(void *q, uint32_t mode, char *items, uint64_t n, uint32_t rounds, int64_t *order)
{
	if (mode == 0)
		hipLaunchKernelGGL(k_queue_enqueue, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dqueue *)q, items, n);
	else if (mode == 1)
		hipLaunchKernelGGL(k_queue_shuffle, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dqueue *)q, items, n,
				   rounds);
	else
		hipLaunchKernelGGL(k_queue_serial, dim3(1), dim3(64), 0, 0, (kma_dqueue *)q, items, n, order);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ---------------- linked pointer queue ---------------- */
__global__ void k_lqueue_init(kma_dlqueue *q)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		q->magic = KMA_LQUEUE_MAGIC;
		q->stub = 0;
		q->head = (uint64_t)(uintptr_t)&q->stub;
		q->tail = (uint64_t)(uintptr_t)&q->stub;
	}
}

__global__ void k_lqueue_enqueue(kma_dlqueue *q, char *items, uint64_t n, unsigned long long *lost)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && !kma_lq_enqueue(q, items + 16u * i))
		atomicAdd(lost, 1ull);
}

__global__ void k_lqueue_shuffle(kma_dlqueue *q, char *items, uint64_t n, uint32_t rounds, unsigned long long *lost)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (!kma_lq_enqueue(q, items + 16u * i))
		atomicAdd(lost, 1ull);
	for (uint32_t r = 0; r < rounds; r++) {
		/* every lane calls (a NULL item is a no-op), as the wavefront
		 * aggregation serves the lanes that call together */
		void *it = kma_lq_dequeue(q);
		if (!kma_lq_enqueue(q, it) && it)
			atomicAdd(lost, 1ull);
	}
}

__global__ void k_lqueue_serial(kma_dlqueue *q, char *items, uint64_t n, int64_t *order, unsigned long long *lost)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	for (uint64_t i = 0; i < n; i++)
		if (!kma_lq_enqueue(q, items + 16u * i))
			atomicAdd(lost, 1ull);
	uint64_t k = 0;
	for (;;) {
		char *it = (char *)kma_lq_dequeue(q);
		if (!it)
			break;
		if (k < n)
			order[k] = (int64_t)((it - items) / 16);
		k++;
	}
	for (; k < n; k++)
		order[k] = -1;
}

extern "C" int kmad_lqueue_init(void *q)
{
	hipLaunchKernelGGL(k_lqueue_init, dim3(1), dim3(64), 0, 0, (kma_dlqueue *)q);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_lqueue_run(void *q, uint32_t mode, char *items, uint64_t n, uint32_t rounds, int64_t *order,
			       unsigned long long *lost)
{
	if (mode == 0)
		hipLaunchKernelGGL(k_lqueue_enqueue, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dlqueue *)q, items, n,
				   lost);
	else if (mode == 1)
		hipLaunchKernelGGL(k_lqueue_shuffle, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dlqueue *)q, items, n,
				   rounds, lost);
	else
		hipLaunchKernelGGL(k_lqueue_serial, dim3(1), dim3(64), 0, 0, (kma_dlqueue *)q, items, n, order, lost);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ---------------- array list ---------------- */
#define LIST_BLOCK 256

__global__ void __launch_bounds__(LIST_BLOCK) k_list_grow(kma_dlist *l, const uint32_t *objs, uint64_t n,
							   int local, uint64_t *out)
{
	__shared__ kma_list_lds<LIST_BLOCK> L;
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t o = i < n ? objs[i] : 0u;
	char *p;
	if (local)
		p = (char *)kma_list_grow_local<LIST_BLOCK>(l, o, L);
	else
		p = o ? (char *)kma_list_grow(l, o) : nullptr;
	if (i >= n)
		return;
	if (p)
		for (uint32_t k = 0; k < o; k++)
			*(uint32_t *)(p + (uint64_t)k * l->obj_size) = (uint32_t)i ^ k;
	out[i] = p ? (uint64_t)(p - (char *)l->heap) : 0u;
}

__global__ void k_list_get(kma_dlist *l, const uint64_t *idx, uint64_t n, uint64_t *out)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	char *p = (char *)kma_list_get(l, idx[i]);
	out[i] = p ? (uint64_t)(p - (char *)l->heap) : 0u;
}

__global__ void k_list_clear(kma_dlist *l, uint64_t n)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		kma_list_clear(l);
}

extern "C" int kmad_list_grow(void *l, const uint32_t *objs, uint64_t n, uint32_t block, int local, uint64_t *out)
{
	if (block == 0 || block > LIST_BLOCK || (block & 63))
		return -1;
	hipLaunchKernelGGL(k_list_grow, dim3(nblocks(n, block)), dim3(block), 0, 0, (kma_dlist *)l, objs, n, local,
			   out);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_list_get(void *l, const uint64_t *idx, uint64_t n, uint64_t *out)
{
	hipLaunchKernelGGL(k_list_get, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dlist *)l, idx, n, out);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_list_clear(void *l, uint64_t n)
{
	hipLaunchKernelGGL(k_list_clear, dim3(nblocks(n, 256)), dim3(256), 0, 0, (kma_dlist *)l, n);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ---------------- binary search tree ---------------- */
__global__ void k_tree_add(kma_dheap *H, kma_tree *t, const uint32_t *keys, uint64_t n, uint32_t *added)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	kma_tree_node *nd = (kma_tree_node *)kma_malloc(H, sizeof(kma_tree_node));
	if (!nd) {
		added[i] = 2u; /* the heap could not serve the node */
		return;
	}
	nd->key = keys[i];
	uint32_t r = kma_tree_add(t, nd);
	if (!r)
		kma_free(H, nd);
	added[i] = r;
}

__global__ void k_tree_get(kma_tree *t, const uint32_t *query, uint64_t n, uint32_t *found)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	kma_tree_node *nd = kma_tree_get(t, query[i]);
	found[i] = nd && nd->key == query[i];
}

/* In-order walk with an explicit stack (cap entries); optionally frees each
 * node after its right child has been read. */
__global__ void k_tree_walk(kma_dheap *H, kma_tree *t, kma_tree_node **stack, uint64_t cap, uint32_t *inorder,
			    uint64_t *res, int free_nodes)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	uint64_t sp = 0, n = 0, bad = 0, overflow = 0;
	uint32_t prev = 0;
	kma_tree_node *c = t->root;
	while (c || sp) {
		while (c) {
			if (sp == cap) {
				overflow = 1;
				break;
			}
			stack[sp++] = c;
			c = c->left;
		}
		if (overflow)
			break;
		c = stack[--sp];
		if (n && c->key <= prev)
			bad++;
		prev = c->key;
		if (n < cap)
			inorder[n] = c->key;
		n++;
		kma_tree_node *r = c->right;
		if (free_nodes)
			kma_free(H, c);
		c = r;
	}
	if (free_nodes && !overflow)
		t->root = nullptr;
	res[0] = n;
	res[1] = bad + overflow;
}

extern "C" int kmad_tree_run(void *heap, void *tree, const uint32_t *keys, uint64_t n, uint32_t *added,
			     const uint32_t *query, uint64_t nq, uint32_t *found, void *stack, uint32_t *inorder,
			     uint64_t *res)
{
	kma_dheap *H = (kma_dheap *)heap;
	kma_tree *t = (kma_tree *)tree;
	if (n)
		hipLaunchKernelGGL(k_tree_add, dim3(nblocks(n, 256)), dim3(256), 0, 0, H, t, keys, n, added);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	if (nq)
		hipLaunchKernelGGL(k_tree_get, dim3(nblocks(nq, 256)), dim3(256), 0, 0, t, query, nq, found);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	hipLaunchKernelGGL(k_tree_walk, dim3(1), dim3(64), 0, 0, H, t, (kma_tree_node **)stack, n + 1, inorder, res, 1);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}
