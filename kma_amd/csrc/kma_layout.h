/*
 * kma_layout.h - native heap layout computation (pure C, no device calls).
 * The header struct mirrors kma_dheap in include/kma_device.h byte for byte.
 */
#ifndef KMA_LAYOUT_H
#define KMA_LAYOUT_H

#include <stddef.h>
#include <stdint.h>

#include "kma.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nat_hdr {
	uint64_t magic;
	uint64_t bytes;
	uint64_t nsb;
	uint64_t npages;
	uint32_t sb_log2;
	uint32_t page_log2;
	uint32_t nclasses;
	uint32_t nctx;
	uint32_t npool;
	uint32_t bm_words;
	uint32_t max_small;
	uint32_t pad0;
	uint64_t off_state;
	uint64_t off_next;
	uint64_t off_bitmap;
	uint64_t off_heads;
	uint64_t off_pool;
	uint64_t off_pbitmap;
	uint64_t off_plen;
	uint64_t off_data;
	uint64_t off_pages;
	uint64_t off_stats;
	uint32_t class_size[32];
	uint32_t class_cap[32];
	uint64_t class_magic[32];
	uint8_t class_for_log2[64];
	uint64_t off_psum;
	uint64_t off_spare;
	uint32_t spare_slots;
	uint32_t pad1;
	uint64_t off_pidx;
	uint32_t pidx_words;
	uint32_t no_partial;
} nat_hdr;

#define NAT_MAGIC 0x4e414d4b31414d4bull
#define NAT_HEADS_PER_CTX 32u
#define NAT_POOL_STRIDE_BYTES 128u
#define NAT_STATS_BYTES (256u + 128u * 8u) /* counters + one line per XCD (shard mask, clean request) */
#define NAT_PAGE_LEVELS 7u

/* Validate `in`, fill defaults into *out and the header (metadata offsets;
 * off_data holds the metadata size until kma_native_place runs).
 * Returns 0, or -1 with a message in err. */
__attribute__((visibility("hidden"))) int kma_native_layout(const kma_config *in, kma_config *out, nat_hdr *nh,
						  char *err, size_t errlen);

/* Bytes to allocate for the heap (metadata + alignment slack + data). */
__attribute__((visibility("hidden"))) uint64_t kma_native_alloc_bytes(const nat_hdr *nh);

/* Place superblock data and pages for a buffer at device address `base`,
 * so both are aligned absolutely to their sizes. */
__attribute__((visibility("hidden"))) void kma_native_place(nat_hdr *nh, uint64_t base, uint64_t total);

#ifdef __cplusplus
}
#endif
#endif
