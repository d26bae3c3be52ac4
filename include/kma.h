/*
 * kma.h - host C API of the MI355X device allocator (libkma.so).
 *
 * Plain C ABI: no HIP, OpenCL or torch types in any signature.  Every entry
 * point blocks until its device work is complete, as the reference's do
 * (kma.c:50,119 call clFinish).
 */
#ifndef KMA_H
#define KMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Heap layouts.
 * KMA_LAYOUT_REFERENCE: the reference's byte layout (kma.h:71-85), served by
 *   kma_malloc(struct clheap *, size_t) in kma_ref_device.h.  At most 65,536
 *   superblocks (the reference's 16-bit queue index, clIndexedQueue.cl:89).
 * KMA_LAYOUT_NATIVE: this library's gfx950 layout, served by
 *   kma_malloc(kma_dheap *, size_t) in kma_device.h. */
#define KMA_LAYOUT_REFERENCE 0u
#define KMA_LAYOUT_NATIVE    1u

#define KMA_MAX_CLASSES 24u

typedef struct kma_config {
	uint32_t layout;      /* KMA_LAYOUT_* */
	uint32_t sb_log2;     /* superblock size = 1 << sb_log2 (reference: 12) */
	uint64_t sblocks;     /* number of superblocks */
	/* NATIVE only (0 selects the default) */
	uint32_t nctx;        /* allocation contexts (power of two, 8..65536; 0: the largest keeping one
				 superblock per class and context within a quarter of the heap, <= 4096) */
	uint32_t npool;       /* superblock pool shards (power of two, 16..512; 0: min(2 * nctx, 512)) */
	uint32_t nclasses;    /* entries used in class_size[] */
	uint32_t class_size[KMA_MAX_CLASSES]; /* ascending block sizes, multiples of 8 */
	uint32_t page_log2;   /* large-request page size log2 (default = sb_log2) */
	uint64_t npages;      /* pages in the large-request region (0 = none) */
	uint32_t no_partial;  /* 1: never reattach partly free superblocks (no partial index, no pool-pressure
				 routing): a detached superblock is reused only once empty, as in the reference */
} kma_config;

typedef struct kma_heap kma_heap; /* opaque host handle */

/* kma_create(dev, ctx, cq, prg, sblocks), kma.h:48-49 / kma.c:79-127:
 * reference layout, 4 KiB superblocks, `sblocks` superblocks; NULL for
 * sblocks == 0 or on any device error (kma.c:89-90, 105-108). */
kma_heap *kma_create(unsigned int sblocks);

/* Extended constructor: layout, superblock size, class table, contexts,
 * pool shards and page region.  NULL on a rejected configuration. */
kma_heap *kma_create_ex(const kma_config *cfg);

/* The reference has no destroy; callers release the cl_mem themselves
 * (tb_kma.c:316).  This frees the device buffer and the handle. */
void kma_destroy(kma_heap *h);

/* Device address to pass to kernels: `struct clheap *` for the reference
 * layout, `kma_dheap *` for the native layout. */
void *kma_device_heap(const kma_heap *h);

/* Bytes of the device buffer (reference layout: sblocks*SB + header,
 * kma.c:37). */
uint64_t kma_heap_bytes(const kma_heap *h);

/* Configuration the heap was built with (defaults filled in). */
int kma_get_config(const kma_heap *h, kma_config *out);

/* Copy `bytes` bytes of the heap image, starting at `offset`, to host memory. */
int kma_snapshot(const kma_heap *h, uint64_t offset, void *dst, uint64_t bytes);

/* Re-run heap initialisation (all blocks free, metadata at its initial image). */
int kma_reset(kma_heap *h);

/* Last error message of the calling thread ("" if none). */
const char *kma_last_error(void);

/* 0 when a gfx9 device is usable, negative otherwise. */
int kma_device_ok(void);

#ifdef __cplusplus
}
#endif
#endif /* KMA_H */
