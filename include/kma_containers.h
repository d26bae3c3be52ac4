/*
 * kma_containers.h - host C API of the data structures built on the device
 * allocator: the bump allocator (reference pma.*), the pointer queue
 * (clQueue.*), the array list (clArrayList.*) and the binary search tree
 * (clTree.*), with the test runners that exercise their device sides.
 * Plain C ABI; every call blocks until its device work is complete.
 */
#ifndef KMA_CONTAINERS_H
#define KMA_CONTAINERS_H

#include "kma.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- bump allocator ---- */
typedef struct kma_pma kma_pma;

/* pma_create(dev, ctx, cq, prg, bytes), pma.h:45-46 / pma.c:79-127:
 * `bytes` of data (the header is extra), sizes rounded up to `align` (a
 * power of two; 1 = unaligned like the reference).  NULL for bytes == 0
 * (pma.c:89-90). */
kma_pma *kma_pma_create(uint64_t bytes, uint32_t align);
void kma_pma_destroy(kma_pma *p);
/* Device address to pass to kernels (kma_pma_dheap *, kma_pma.h). */
void *kma_pma_device(const kma_pma *p);
/* Head back to 0 (clheap_init, pma.cl:24-30). */
int kma_pma_reset(kma_pma *p);
/* Current head (bytes handed out plus failed requests' sizes). */
uint64_t kma_pma_head(const kma_pma *p);
/* Test runner: n requests of sizes[i] bytes.  serial = 1: one work-item
 * issues them in order; serial = 0: work-item i issues request i (wavefronts
 * aggregate).  out[i] = data offset of the block, or UINT64_MAX for NULL. */
int kma_pma_run(kma_pma *p, const uint32_t *sizes, uint64_t n, int serial, uint64_t *out);

/* ---- pointer queue ---- */
typedef struct kma_queue kma_queue;

/* clQueue_create(dev, ctx, cq, prg), clQueue.h:68: an empty queue.  The ring
 * holds at most `capacity` items (rounded up to a power of two). */
kma_queue *kma_queue_create(uint64_t capacity);
void kma_queue_destroy(kma_queue *q);
/* Device address to pass to kernels (kma_dqueue *, kma_queue.h). */
void *kma_queue_device(const kma_queue *q);
/* Empty the queue (clQueue_init, clQueue.cl:27-32). */
int kma_queue_reset(kma_queue *q);
/* Host view of the queue: the items from head to tail, oldest first, into
 * items[0..max).  Returns the number of items, or -1 if a slot between head
 * and tail is not filled (only possible while device calls are running). */
int64_t kma_queue_items(const kma_queue *q, uint64_t *items, uint64_t max);
/* Test runner on a runner-owned item array (item i = base + 16 i; *base_out
 * receives base):
 *   mode 0 clQueue_test_enqueue (clQueue.cl:121-137): work-item i enqueues
 *          item i;
 *   mode 1 clQueue_test_dequeue (clQueue.cl:140-162): work-item i enqueues
 *          item i, then `rounds` times dequeues one item and, if it got
 *          one, enqueues it again;
 *   mode 2 one work-item enqueues items 0..nwi-1, then dequeues until the
 *          queue is empty; order_out[k] = index of the k-th item dequeued
 *          (-1 past the end). */
int kma_queue_run(kma_queue *q, uint32_t mode, uint64_t nwi, uint32_t rounds, uint64_t *base_out,
		  int64_t *order_out);

/* ---- unbounded linked pointer queue ---- */
typedef struct kma_lqueue kma_lqueue;

/* clQueue_create(dev, ctx, cq, prg), clQueue.h:68, with the reference's
 * contract: items are linked through their first 8 bytes
 * (clqueue_item.next, clQueue.h:42-45), so the queue holds any number of
 * them and enqueue never fails for capacity (kma_lqueue.h). */
kma_lqueue *kma_lqueue_create(void);
void kma_lqueue_destroy(kma_lqueue *q);
/* Device address to pass to kernels (kma_dlqueue *, kma_lqueue.h). */
void *kma_lqueue_device(const kma_lqueue *q);
/* Empty the queue (clQueue_init, clQueue.cl:27-32). */
int kma_lqueue_reset(kma_lqueue *q);
/* The reference's validator (tb_clQueue.c:126-157) on the chain from head:
 * it must end (no cycle), its last node must be the one tail names, and
 * every node must be an item of the runner's array or the queue's own stub,
 * each at most once. */
typedef struct kma_lqueue_report {
	uint64_t items;    /* items on the chain (the stub not counted) */
	uint64_t tail_ok;  /* 1: the chain ends at the node tail names */
	uint64_t cycle;    /* 1: the chain did not end within items + 1 nodes */
	uint64_t stubs;    /* times the chain passes the stub (0 or 1) */
	uint64_t outside;  /* links to neither an item nor the stub */
	uint64_t repeats;  /* items met twice */
	uint64_t lost;     /* enqueue or dequeue calls that gave up (corrupt queue) */
} kma_lqueue_report;
/* Test runner on a runner-owned item array of nwi 16-byte items (the modes
 * of kma_queue_run): 0 every work-item enqueues its item; 1 then `rounds`
 * times dequeues one item and enqueues it again; 2 one work-item enqueues
 * items 0..nwi-1 and dequeues until empty, order_out[k] = index of the k-th
 * item dequeued (-1 past the end).  The chain is validated after the
 * kernels (*rep). */
int kma_lqueue_run(kma_lqueue *q, uint32_t mode, uint64_t nwi, uint32_t rounds, int64_t *order_out,
		   kma_lqueue_report *rep);

/* ---- array list ---- */
typedef struct kma_list kma_list;

/* clArrayList_create(dev, ctx, cq, prg, objsize, heap), clArrayList.h:48-49:
 * an empty list of obj_size-byte objects whose pages come from `heap` (a
 * native-layout heap), at most max_pages pages. */
kma_list *kma_list_create(kma_heap *heap, uint32_t obj_size, uint64_t max_pages);
void kma_list_destroy(kma_list *l);
/* Device address to pass to kernels (kma_dlist *, kma_list.h). */
void *kma_list_device(const kma_list *l);
/* Test runner: work-item i asks for objs[i] objects, in work-groups of
 * `block` threads (multiple of 64, <= 256); local = 1: kma_list_grow_local
 * (one page per work-group), local = 0: kma_list_grow per work-item.  Each
 * work-item writes every object it got (u32 pattern i ^ k at each object's
 * start) and out[i] = heap offset of its first object (0 for NULL). */
int kma_list_run_grow(kma_list *l, const uint32_t *objs, uint64_t nwi, uint32_t block, int local,
		      uint64_t *out);
/* out[k] = heap offset of object idx[k] (kma_list_get), 0 past the end. */
int kma_list_run_get(kma_list *l, const uint64_t *idx, uint64_t n, uint64_t *out);
/* clArrayList_clear (clArrayList.cl:206-216) from `nwi` work-items. */
int kma_list_clear(kma_list *l, uint64_t nwi);
/* Heap offsets of the pages in queue order; returns their number (or -1). */
int64_t kma_list_pages(const kma_list *l, uint64_t *pages, uint64_t max);

/* ---- binary search tree ---- */
/* Test runner for kma_tree_add / kma_tree_get (clTree.cl:27-74) over nodes
 * taken from `heap` (native layout): work-item i mallocs a node with key
 * keys[i] and adds it (added[i] = 1) or frees it again (added[i] = 0, key
 * already present); then work-item j looks up query[j] (found[j] = 1 if a
 * node with that key is found and carries it).  One work-item then walks
 * the tree in order: inorder[0..*ninorder) receives the keys and
 * *violations counts keys not strictly above their predecessor.  Finally
 * every node is freed again. */
int kma_tree_run(kma_heap *heap, const uint32_t *keys, uint64_t n, uint32_t *added, const uint32_t *query,
		 uint64_t nq, uint32_t *found, uint32_t *inorder, uint64_t *ninorder, uint64_t *violations);

#ifdef __cplusplus
}
#endif
#endif /* KMA_CONTAINERS_H */
