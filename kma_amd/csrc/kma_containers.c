/*
 * kma_containers.c - host side of the bump allocator, pointer queue, array
 * list and binary search tree (include/kma_containers.h).  C over the HIP
 * runtime C API; the kernels are in kma_containers.hip.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_containers.h"
#include "kma_internal.h"
#include "kma_launch.h"

/* Mirrors of the device headers' fixed layouts (kma_pma.h, kma_queue.h). */
#define PMA_HDR 256u
#define PMA_HEAD_OFF 128u
#define QUEUE_HDR 384u
#define QUEUE_HEAD_OFF 128u
#define QUEUE_TAIL_OFF 256u
#define LIST_PAGE_HDR 24u

struct kma_pma {
	char *dev;
	uint64_t bytes;
	uint32_t align;
};

struct kma_queue {
	char *dev;
	uint64_t cap;
};

struct kma_list {
	char *dev; /* kma_dlist */
	kma_heap *heap;
	kma_queue *queue;
	uint32_t obj_size;
};

/* Device-side list header: {magic, obj_size, heap, queue} (kma_list.h). */
#define LIST_MAGIC 0x314c414d4b4d414bull

/* device allocation with the error path as a statement */
#define DEV_ALLOC(ptr, bytes, ret)                                                   \
	do {                                                                         \
		if (hipMalloc((void **)&(ptr), (bytes) > 0 ? (bytes) : 1) != hipSuccess) { \
			set_err("device allocation of %llu bytes failed",            \
				(unsigned long long)(bytes));                        \
			ret;                                                         \
		}                                                                    \
	} while (0)

/* ---------------- bump allocator ---------------- */
kma_pma *kma_pma_create(uint64_t bytes, uint32_t align)
{
	if (bytes == 0) { /* pma.c:89-90 */
		set_err("bytes == 0");
		return NULL;
	}
	if (align == 0 || (align & (align - 1))) {
		set_err("align must be a power of two");
		return NULL;
	}
	if (kma_device_ok() != 0)
		return NULL;
	kma_pma *p = (kma_pma *)calloc(1, sizeof(*p));
	if (!p)
		return NULL;
	p->bytes = bytes;
	p->align = align;
	if (hipMalloc((void **)&p->dev, PMA_HDR + bytes) != hipSuccess) {
		set_err("PMA: Could not allocate heap on-device (%llu bytes)", (unsigned long long)bytes);
		free(p);
		return NULL;
	}
	if (kma_pma_reset(p) != 0) {
		hipFree(p->dev);
		free(p);
		return NULL;
	}
	return p;
}

void kma_pma_destroy(kma_pma *p)
{
	if (!p)
		return;
	hipFree(p->dev);
	free(p);
}

void *kma_pma_device(const kma_pma *p)
{
	return p ? p->dev : NULL;
}

int kma_pma_reset(kma_pma *p)
{
	if (!p)
		return -1;
	HIPCHK(hipMemset(p->dev, 0, PMA_HDR), -1);
	LCHK(kmad_pma_init(p->dev, p->bytes, p->align), -1);
	return 0;
}

uint64_t kma_pma_head(const kma_pma *p)
{
	uint64_t h = 0;
	if (p)
		hipMemcpy(&h, p->dev + PMA_HEAD_OFF, 8, hipMemcpyDeviceToHost);
	return h;
}

int kma_pma_run(kma_pma *p, const uint32_t *sizes, uint64_t n, int serial, uint64_t *out)
{
	if (!p || (n && (!sizes || !out))) {
		set_err("bad arguments");
		return -1;
	}
	uint32_t *ds = NULL;
	uint64_t *dout = NULL;
	int rc = -1;
	DEV_ALLOC(ds, n * 4, goto done);
	DEV_ALLOC(dout, n * 8, goto done);
	if (hipMemcpy(ds, sizes, n * 4, hipMemcpyHostToDevice) != hipSuccess) {
		set_err("upload failed");
		goto done;
	}
	if (kmad_pma_run(p->dev, ds, n, serial, dout) != 0 ||
	    hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("PMA run failed");
		goto done;
	}
	rc = 0;
done:
	if (ds)
		hipFree(ds);
	if (dout)
		hipFree(dout);
	return rc;
}

/* ---------------- pointer queue ---------------- */
kma_queue *kma_queue_create(uint64_t capacity)
{
	if (capacity == 0 || capacity > (1ull << 32)) {
		set_err("queue capacity must be in 1..2^32");
		return NULL;
	}
	if (kma_device_ok() != 0)
		return NULL;
	uint64_t cap = 1;
	while (cap < capacity)
		cap <<= 1;
	kma_queue *q = (kma_queue *)calloc(1, sizeof(*q));
	if (!q)
		return NULL;
	q->cap = cap;
	if (hipMalloc((void **)&q->dev, QUEUE_HDR + cap * 16) != hipSuccess) {
		set_err("queue: device allocation failed");
		free(q);
		return NULL;
	}
	if (kma_queue_reset(q) != 0) {
		hipFree(q->dev);
		free(q);
		return NULL;
	}
	return q;
}

void kma_queue_destroy(kma_queue *q)
{
	if (!q)
		return;
	hipFree(q->dev);
	free(q);
}

void *kma_queue_device(const kma_queue *q)
{
	return q ? q->dev : NULL;
}

int kma_queue_reset(kma_queue *q)
{
	if (!q)
		return -1;
	HIPCHK(hipMemset(q->dev, 0, QUEUE_HDR), -1);
	LCHK(kmad_queue_init(q->dev, q->cap), -1);
	return 0;
}

int64_t kma_queue_items(const kma_queue *q, uint64_t *items, uint64_t max)
{
	if (!q)
		return -1;
	uint64_t ht[2];
	HIPCHK(hipMemcpy(&ht[0], q->dev + QUEUE_HEAD_OFF, 8, hipMemcpyDeviceToHost), -1);
	HIPCHK(hipMemcpy(&ht[1], q->dev + QUEUE_TAIL_OFF, 8, hipMemcpyDeviceToHost), -1);
	uint64_t n = ht[1] - ht[0];
	if (n > q->cap) {
		set_err("queue head/tail inconsistent");
		return -1;
	}
	uint64_t *sl = (uint64_t *)malloc((q->cap ? q->cap : 1) * 16);
	if (!sl)
		return -1;
	if (hipMemcpy(sl, q->dev + QUEUE_HDR, q->cap * 16, hipMemcpyDeviceToHost) != hipSuccess) {
		free(sl);
		set_err("queue download failed");
		return -1;
	}
	int64_t rc = (int64_t)n;
	for (uint64_t k = 0; k < n; k++) {
		uint64_t t = ht[0] + k, i = t & (q->cap - 1);
		if (sl[2 * i] != t + 1) {
			set_err("slot of ticket %llu not filled", (unsigned long long)t);
			rc = -1;
			break;
		}
		if (items && k < max)
			items[k] = sl[2 * i + 1];
	}
	free(sl);
	return rc;
}

int kma_queue_run(kma_queue *q, uint32_t mode, uint64_t nwi, uint32_t rounds, uint64_t *base_out, int64_t *order_out)
{
	if (!q || mode > 2 || nwi > q->cap || (mode == 2 && !order_out)) {
		set_err("bad queue run (mode %u, %llu work-items, capacity %llu)", mode, (unsigned long long)nwi,
			(unsigned long long)(q ? q->cap : 0));
		return -1;
	}
	char *items = NULL;
	int64_t *dorder = NULL;
	int rc = -1;
	DEV_ALLOC(items, nwi * 16, goto done);
	if (mode == 2)
		DEV_ALLOC(dorder, nwi * 8, goto done);
	if (base_out)
		*base_out = (uint64_t)(uintptr_t)items;
	if (kmad_queue_run(q->dev, mode, items, nwi, rounds, dorder) != 0) {
		set_err("queue kernels failed");
		goto done;
	}
	if (mode == 2 && hipMemcpy(order_out, dorder, nwi * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("download failed");
		goto done;
	}
	rc = 0;
done:
	/* the queue keeps the items' addresses; callers compare them as values
	 * against *base_out, nothing dereferences them after this run */
	if (items)
		hipFree(items);
	if (dorder)
		hipFree(dorder);
	return rc;
}

/* ---------------- linked pointer queue ---------------- */
/* header words (kma_dlqueue in kma_lqueue.h) */
#define LQ_BYTES 512u
#define LQ_HEAD 16u
#define LQ_TAIL 32u
#define LQ_STUB 48u
#define LQ_PTR ((1ull << 48) - 1ull)

struct kma_lqueue {
	char *dev;
};

kma_lqueue *kma_lqueue_create(void)
{
	if (kma_device_ok() != 0)
		return NULL;
	kma_lqueue *q = (kma_lqueue *)calloc(1, sizeof(*q));
	if (!q)
		return NULL;
	if (hipMalloc((void **)&q->dev, LQ_BYTES) != hipSuccess) {
		set_err("linked queue: device allocation failed");
		free(q);
		return NULL;
	}
	if (((uint64_t)(uintptr_t)q->dev + LQ_BYTES) & ~LQ_PTR) {
		set_err("linked queue: device address above 2^48");
		hipFree(q->dev);
		free(q);
		return NULL;
	}
	if (kma_lqueue_reset(q) != 0) {
		hipFree(q->dev);
		free(q);
		return NULL;
	}
	return q;
}

void kma_lqueue_destroy(kma_lqueue *q)
{
	if (!q)
		return;
	hipFree(q->dev);
	free(q);
}

void *kma_lqueue_device(const kma_lqueue *q)
{
	return q ? q->dev : NULL;
}

int kma_lqueue_reset(kma_lqueue *q)
{
	if (!q)
		return -1;
	HIPCHK(hipMemset(q->dev, 0, LQ_BYTES), -1);
	LCHK(kmad_lqueue_init(q->dev), -1);
	return 0;
}

/* Walk the chain from head over host copies of the header and the items. */
static void lqueue_walk(const uint64_t *hdr, uint64_t dev_hdr, const uint64_t *items, uint64_t dev_items,
			uint64_t n, kma_lqueue_report *r)
{
	uint8_t *seen = (uint8_t *)calloc(n ? n : 1, 1);
	const uint64_t stub = dev_hdr + 8u * LQ_STUB, tail = hdr[LQ_TAIL] & LQ_PTR;
	uint64_t cur = hdr[LQ_HEAD] & LQ_PTR, last = 0, steps = 0;
	while (cur && steps <= n + 1u) {
		uint64_t link;
		if (cur == stub) {
			r->stubs++;
			link = hdr[LQ_STUB];
		} else if (cur >= dev_items && cur < dev_items + 16u * n && (cur - dev_items) % 16u == 0) {
			uint64_t i = (cur - dev_items) / 16u;
			if (seen && seen[i]++)
				r->repeats++;
			r->items++;
			link = items[2u * i];
		} else {
			r->outside++;
			break;
		}
		last = cur;
		cur = link & LQ_PTR;
		steps++;
	}
	r->cycle = cur != 0 && r->outside == 0;
	r->tail_ok = !r->cycle && !r->outside && last == tail;
	free(seen);
}

int kma_lqueue_run(kma_lqueue *q, uint32_t mode, uint64_t nwi, uint32_t rounds, int64_t *order_out,
		   kma_lqueue_report *rep)
{
	if (!q || mode > 2 || (mode == 2 && !order_out) || !rep) {
		set_err("bad linked-queue run (mode %u)", mode);
		return -1;
	}
	memset(rep, 0, sizeof(*rep));
	char *items = NULL;
	int64_t *dorder = NULL;
	unsigned long long *dlost = NULL, lost = 0;
	uint64_t *hitems = NULL, hdr[LQ_BYTES / 8];
	int rc = -1;
	DEV_ALLOC(items, (nwi ? nwi : 1) * 16, goto done);
	DEV_ALLOC(dlost, 8, goto done);
	if (mode == 2)
		DEV_ALLOC(dorder, (nwi ? nwi : 1) * 8, goto done);
	if (((uint64_t)(uintptr_t)items + nwi * 16) & ~LQ_PTR) {
		set_err("linked queue: item address above 2^48");
		goto done;
	}
	if (hipMemset(dlost, 0, 8) != hipSuccess ||
	    kmad_lqueue_run(q->dev, mode, items, nwi, rounds, dorder, dlost) != 0) {
		set_err("linked-queue kernels failed");
		goto done;
	}
	hitems = (uint64_t *)malloc((nwi ? nwi : 1) * 16);
	if (!hitems || hipMemcpy(hitems, items, nwi * 16, hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(hdr, q->dev, LQ_BYTES, hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(&lost, dlost, 8, hipMemcpyDeviceToHost) != hipSuccess ||
	    (mode == 2 && hipMemcpy(order_out, dorder, nwi * 8, hipMemcpyDeviceToHost) != hipSuccess)) {
		set_err("download failed");
		goto done;
	}
	lqueue_walk(hdr, (uint64_t)(uintptr_t)q->dev, hitems, (uint64_t)(uintptr_t)items, nwi, rep);
	rep->lost = lost;
	rc = 0;
done:
	/* the items leave with this run: empty the queue so that no link to
	 * them survives */
	if (items && kma_lqueue_reset(q) != 0)
		rc = -1;
	free(hitems);
	if (items)
		hipFree(items);
	if (dlost)
		hipFree(dlost);
	if (dorder)
		hipFree(dorder);
	return rc;
}

/* ---------------- array list ---------------- */
kma_list *kma_list_create(kma_heap *heap, uint32_t obj_size, uint64_t max_pages)
{
	if (!heap || heap->cfg.layout != KMA_LAYOUT_NATIVE || obj_size < 4 || max_pages == 0) {
		set_err("array list needs a native heap, obj_size >= 4 and max_pages >= 1");
		return NULL;
	}
	kma_list *l = (kma_list *)calloc(1, sizeof(*l));
	if (!l)
		return NULL;
	l->heap = heap;
	l->obj_size = obj_size;
	l->queue = kma_queue_create(max_pages);
	if (!l->queue) {
		free(l);
		return NULL;
	}
	uint64_t hdr[4] = {LIST_MAGIC, obj_size, (uint64_t)(uintptr_t)heap->dev,
			   (uint64_t)(uintptr_t)kma_queue_device(l->queue)};
	if (hipMalloc((void **)&l->dev, sizeof(hdr)) != hipSuccess ||
	    hipMemcpy(l->dev, hdr, sizeof(hdr), hipMemcpyHostToDevice) != hipSuccess) {
		set_err("array list: device allocation failed");
		kma_list_destroy(l);
		return NULL;
	}
	return l;
}

void kma_list_destroy(kma_list *l)
{
	if (!l)
		return;
	if (l->dev)
		hipFree(l->dev);
	kma_queue_destroy(l->queue);
	free(l);
}

void *kma_list_device(const kma_list *l)
{
	return l ? l->dev : NULL;
}

int kma_list_run_grow(kma_list *l, const uint32_t *objs, uint64_t nwi, uint32_t block, int local, uint64_t *out)
{
	if (!l || !objs || !out || block == 0 || block > 256 || (block & 63)) {
		set_err("bad arguments");
		return -1;
	}
	uint32_t *dobjs = NULL;
	uint64_t *dout = NULL;
	int rc = -1;
	DEV_ALLOC(dobjs, nwi * 4, goto done);
	DEV_ALLOC(dout, nwi * 8, goto done);
	if (hipMemcpy(dobjs, objs, nwi * 4, hipMemcpyHostToDevice) != hipSuccess ||
	    kmad_list_grow(l->dev, dobjs, nwi, block, local, dout) != 0 ||
	    hipMemcpy(out, dout, nwi * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("array list grow failed");
		goto done;
	}
	rc = 0;
done:
	if (dobjs)
		hipFree(dobjs);
	if (dout)
		hipFree(dout);
	return rc;
}

int kma_list_run_get(kma_list *l, const uint64_t *idx, uint64_t n, uint64_t *out)
{
	if (!l || (n && (!idx || !out))) {
		set_err("bad arguments");
		return -1;
	}
	uint64_t *didx = NULL, *dout = NULL;
	int rc = -1;
	DEV_ALLOC(didx, n * 8, goto done);
	DEV_ALLOC(dout, n * 8, goto done);
	if (hipMemcpy(didx, idx, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
	    kmad_list_get(l->dev, didx, n, dout) != 0 ||
	    hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("array list get failed");
		goto done;
	}
	rc = 0;
done:
	if (didx)
		hipFree(didx);
	if (dout)
		hipFree(dout);
	return rc;
}

int kma_list_clear(kma_list *l, uint64_t nwi)
{
	if (!l || nwi == 0) {
		set_err("bad arguments");
		return -1;
	}
	LCHK(kmad_list_clear(l->dev, nwi), -1);
	return 0;
}

int64_t kma_list_pages(const kma_list *l, uint64_t *pages, uint64_t max)
{
	if (!l)
		return -1;
	int64_t n = kma_queue_items(l->queue, pages, max);
	if (n > 0 && pages)
		for (int64_t k = 0; k < n && (uint64_t)k < max; k++)
			pages[k] -= (uint64_t)(uintptr_t)l->heap->dev;
	return n;
}

/* ---------------- binary search tree ---------------- */
int kma_tree_run(kma_heap *heap, const uint32_t *keys, uint64_t n, uint32_t *added, const uint32_t *query,
		 uint64_t nq, uint32_t *found, uint32_t *inorder, uint64_t *ninorder, uint64_t *violations)
{
	if (!heap || heap->cfg.layout != KMA_LAYOUT_NATIVE || (n && (!keys || !added)) || (nq && (!query || !found)) ||
	    !inorder || !ninorder || !violations) {
		set_err("bad arguments");
		return -1;
	}
	uint32_t *dkeys = NULL, *dadded = NULL, *dq = NULL, *dfound = NULL, *dinorder = NULL;
	uint64_t *droot = NULL, *dres = NULL;
	void *dstack = NULL;
	uint64_t res[2] = {0, 0};
	int rc = -1;
	DEV_ALLOC(dkeys, n * 4, goto done);
	DEV_ALLOC(dadded, n * 4, goto done);
	DEV_ALLOC(dq, nq * 4, goto done);
	DEV_ALLOC(dfound, nq * 4, goto done);
	DEV_ALLOC(dinorder, (n + 1) * 4, goto done);
	DEV_ALLOC(droot, 8, goto done);
	DEV_ALLOC(dres, 16, goto done);
	DEV_ALLOC(dstack, (n + 1) * 8, goto done);
	if (hipMemset(droot, 0, 8) != hipSuccess || (n && hipMemcpy(dkeys, keys, n * 4, hipMemcpyHostToDevice) != hipSuccess) ||
	    (nq && hipMemcpy(dq, query, nq * 4, hipMemcpyHostToDevice) != hipSuccess)) {
		set_err("upload failed");
		goto done;
	}
	if (kmad_tree_run(heap->dev, droot, dkeys, n, dadded, dq, nq, dfound, dstack, dinorder, dres) != 0) {
		set_err("tree kernels failed");
		goto done;
	}
	if (hipMemcpy(res, dres, 16, hipMemcpyDeviceToHost) != hipSuccess ||
	    (n && hipMemcpy(added, dadded, n * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
	    (nq && hipMemcpy(found, dfound, nq * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
	    hipMemcpy(inorder, dinorder, (res[0] < n + 1 ? res[0] : n + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("download failed");
		goto done;
	}
	*ninorder = res[0];
	*violations = res[1];
	rc = 0;
done:
	if (dkeys)
		hipFree(dkeys);
	if (dadded)
		hipFree(dadded);
	if (dq)
		hipFree(dq);
	if (dfound)
		hipFree(dfound);
	if (dinorder)
		hipFree(dinorder);
	if (droot)
		hipFree(droot);
	if (dres)
		hipFree(dres);
	if (dstack)
		hipFree(dstack);
	return rc;
}
