/*
 * kma_detbatch.c - host side of the deterministic batch mode on the
 * reference layout (kma_batch_malloc_det / kma_batch_free_det).  The
 * algorithm is described in kma_detbatch.hip.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_internal.h"
#include "kma_launch.h"
#include "kma_work.h"

#define DCLS 16u

/* device buffers of one call; after the first failed allocation every
 * later one returns NULL too, so checking the last pointer suffices */
typedef struct dbuf {
	void *p[16];
	int n;
	int failed;
} dbuf;

static void *dalloc(dbuf *b, uint64_t bytes)
{
	void *p = NULL;
	if (b->failed || b->n >= 16 || hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
		b->failed = 1;
		return NULL;
	}
	b->p[b->n++] = p;
	return p;
}

static void dfree(dbuf *b)
{
	for (int i = 0; i < b->n; i++)
		hipFree(b->p[i]);
	b->n = 0;
}

static int check_heap(const kma_heap *h)
{
	if (!h || h->cfg.layout != KMA_LAYOUT_REFERENCE || h->cfg.sb_log2 > 17) {
		set_err("the deterministic batch mode runs on a reference-layout heap");
		return -1;
	}
	return 0;
}

int64_t kma_batch_malloc_det(kma_heap *h, const uint32_t *sizes, uint64_t n, uint64_t *out, double *ms)
{
	if (check_heap(h) != 0)
		return -1;
	if (n == 0)
		return 0;
	if (!sizes || !out || n >= 0xffffffffull) {
		set_err("bad arguments");
		return -1;
	}
	uint64_t nb = (n + 255) / 256;
	dbuf b = {{0}, 0, 0};
	hipEvent_t e0 = NULL, e1 = NULL;
	int64_t rc = -1;
	uint32_t *ds = dalloc(&b, n * 4);
	int8_t *cls = dalloc(&b, n);
	uint32_t *rank = dalloc(&b, n * 4);
	uint32_t *cnt = dalloc(&b, nb * DCLS * 4);
	uint32_t *tot = dalloc(&b, DCLS * 4);
	uint32_t *evcnt = dalloc(&b, nb * 4);
	uint32_t *small = dalloc(&b, 64 * 4); /* [0] events [1] pops [2] err, [16..32) coff */
	uint32_t *table = dalloc(&b, (n + DCLS) * 4);
	uint32_t *evc = dalloc(&b, n * 4);
	uint32_t *evq = dalloc(&b, n * 4);
	uint32_t *popped = dalloc(&b, n * 4);
	uint64_t *dout = dalloc(&b, n * 8);
	if (!dout || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		set_err("deterministic batch: device allocation failed");
		goto done;
	}
	if (hipMemcpy(ds, sizes, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(small, 0, 64 * 4) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
		set_err("deterministic batch: upload failed");
		goto done;
	}
	kmad_det_bufs bufs = {cls, rank, cnt, tot, evcnt, small, table, evc, evq, popped};
	hipEventRecord(e0, 0);
	if (kmad_det_malloc(h->dev, ds, n, &bufs, dout, NULL, NULL)) {
		set_err("deterministic batch: launch failed");
		goto done;
	}
	hipEventRecord(e1, 0);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("deterministic batch kernels failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	uint32_t hs[3];
	if (hipMemcpy(hs, small, 12, hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("deterministic batch: download failed");
		goto done;
	}
	if (hs[2]) {
		/* the image is undefined now; callers reset the heap */
		set_err("deterministic batch: %u requests fall in a 1-slot class", hs[2]);
		goto done;
	}
	if (ms) {
		float f = 0;
		hipEventElapsedTime(&f, e0, e1);
		*ms = f;
	}
	rc = 0;
	for (uint64_t i = 0; i < n; i++)
		rc += out[i] == 0;
done:
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	dfree(&b);
	return rc;
}

int kma_batch_malloc_det_noalloc(kma_heap *h, const uint32_t *sizes, uint64_t n, uint64_t *out, uint64_t *status,
				 void *stream)
{
	if (check_heap(h) != 0)
		return -1;
	if (n == 0 || n > KMA_DET_NOALLOC_MAX) {
		set_err("allocation-free batch: n = %llu outside 1..%u (the static scratch holds %u requests)",
			(unsigned long long)n, KMA_DET_NOALLOC_MAX, KMA_DET_NOALLOC_MAX);
		return -1;
	}
	if (!sizes || !out || !status) {
		set_err("bad arguments");
		return -1;
	}
	int rc = kmad_det_malloc_noalloc(h->dev, sizes, n, out, (unsigned long long *)status, stream);
	if (rc != 0) {
		set_err("allocation-free batch: launch failed: %d (%s)", rc, hipGetErrorString((hipError_t)(-rc)));
		return -1;
	}
	return 0;
}

uint64_t kma_det_noalloc_scratch_bytes(void)
{
	return kmad_det_scratch_bytes();
}

int kma_batch_free_det(kma_heap *h, const uint64_t *offs, uint64_t n, double *ms)
{
	if (check_heap(h) != 0)
		return -1;
	if (n == 0)
		return 0;
	if (!offs || n >= 0xffffffffull) {
		set_err("bad arguments");
		return -1;
	}
	uint64_t nb = (n + 255) / 256, S = h->cfg.sblocks;
	dbuf b = {{0}, 0, 0};
	hipEvent_t e0 = NULL, e1 = NULL;
	int rc = -1;
	uint64_t *doffs = dalloc(&b, n * 8);
	uint32_t *minblk = dalloc(&b, S * 4);
	uint32_t *evcnt = dalloc(&b, nb * 4);
	uint32_t *last_sb = dalloc(&b, n * 4);
	uint32_t *order = dalloc(&b, S * 4);
	uint32_t *ne = dalloc(&b, 4);
	if (!ne || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		set_err("deterministic batch free: device allocation failed");
		goto done;
	}
	if (hipMemcpy(doffs, offs, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(minblk, 0xff, S * 4) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
		set_err("deterministic batch free: upload failed");
		goto done;
	}
	hipEventRecord(e0, 0);
	if (kmad_detf_min(h->dev, doffs, n, minblk) ||
	    kmad_detf_words(h->dev, doffs, n, minblk, evcnt, last_sb) || kmad_det_scan(evcnt, nb, 1, ne) ||
	    kmad_detf_order(h->dev, doffs, n, minblk, evcnt, last_sb, order) ||
	    kmad_detf_enqueue(h->dev, order, ne)) {
		set_err("deterministic batch free: launch failed");
		goto done;
	}
	hipEventRecord(e1, 0);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("deterministic batch free kernels failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	if (ms) {
		float f = 0;
		hipEventElapsedTime(&f, e0, e1);
		*ms = f;
	}
	rc = 0;
done:
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	dfree(&b);
	return rc;
}
