/*
 * kma_layout.c - native heap layout computation (pure C, no device calls).
 */
#include "kma_layout.h"

#include <stdio.h>
#include <string.h>

static uint64_t align_up(uint64_t x, uint64_t a)
{
	return (x + a - 1) & ~(a - 1);
}

static int is_pow2(uint64_t x)
{
	return x && !(x & (x - 1));
}

int kma_native_layout(const kma_config *in, kma_config *cfg, nat_hdr *nh, char *err, size_t errlen)
{
	*cfg = *in;
	if (cfg->sb_log2 == 0)
		cfg->sb_log2 = 14;
	if (cfg->sb_log2 < 10 || cfg->sb_log2 > 20) {
		snprintf(err, errlen, "native layout: sb_log2 %u outside 10..20", cfg->sb_log2);
		return -1;
	}
	uint32_t sbsz = 1u << cfg->sb_log2;
	if (cfg->nclasses == 0) {
		uint32_t n = 0;
		for (uint32_t s = 8; s <= sbsz / 2 && n < KMA_MAX_CLASSES; s <<= 1)
			cfg->class_size[n++] = s;
		cfg->nclasses = n;
	}
	if (cfg->nclasses > KMA_MAX_CLASSES) {
		snprintf(err, errlen, "native layout: %u classes > %u", cfg->nclasses, KMA_MAX_CLASSES);
		return -1;
	}
	for (uint32_t c = 0; c < cfg->nclasses; c++) {
		uint32_t s = cfg->class_size[c];
		if (s < 8 || (s & 7) || s > sbsz || (c && s <= cfg->class_size[c - 1])) {
			snprintf(err, errlen, "native layout: bad class table entry %u (%u B)", c, s);
			return -1;
		}
	}
	if (cfg->sblocks == 0 && cfg->npages == 0) {
		snprintf(err, errlen, "native layout: empty heap");
		return -1;
	}
	if (cfg->no_partial > 1) {
		snprintf(err, errlen, "native layout: no_partial %u is not 0 or 1", cfg->no_partial);
		return -1;
	}
	if (cfg->sblocks > 0xfffffff0ull) {
		snprintf(err, errlen, "native layout: too many superblocks");
		return -1;
	}
	/* contexts and pool shards are split evenly over the 8 XCDs; each XCD
	 * has at least one dirty and one clean shard.  Default contexts: the
	 * largest power of two that keeps one superblock per class and context
	 * within a quarter of the heap, at most 4,096 (one per wavefront that
	 * is resident at once at the allocator kernels' register use) */
	if (cfg->nctx == 0) {
		uint64_t want = cfg->sblocks / (4ull * cfg->nclasses);
		uint32_t n = 8;
		while ((uint64_t)n * 2 <= want && n < 4096)
			n *= 2;
		cfg->nctx = n;
	}
	if (!is_pow2(cfg->nctx) || cfg->nctx < 8 || cfg->nctx > 65536) {
		snprintf(err, errlen, "native layout: nctx %u must be a power of two in 8..65536", cfg->nctx);
		return -1;
	}
	if (cfg->npool == 0)
		cfg->npool = cfg->nctx * 2 < 512 ? cfg->nctx * 2 : 512;
	if (!is_pow2(cfg->npool) || cfg->npool < 16 || cfg->npool > 512) {
		snprintf(err, errlen, "native layout: npool %u must be a power of two in 16..512", cfg->npool);
		return -1;
	}
	if (cfg->page_log2 == 0)
		cfg->page_log2 = cfg->sb_log2;
	if (cfg->page_log2 < 12 || cfg->page_log2 > 30) {
		snprintf(err, errlen, "native layout: page_log2 %u outside 12..30", cfg->page_log2);
		return -1;
	}

	memset(nh, 0, sizeof(*nh));
	nh->magic = NAT_MAGIC;
	nh->nsb = cfg->sblocks;
	nh->npages = cfg->npages;
	nh->no_partial = cfg->no_partial;
	nh->sb_log2 = cfg->sb_log2;
	nh->page_log2 = cfg->page_log2;
	nh->nclasses = cfg->nclasses;
	nh->nctx = cfg->nctx;
	nh->npool = cfg->npool;
	uint32_t maxcap = sbsz / cfg->class_size[0];
	nh->bm_words = (maxcap + 31) / 32;
	nh->max_small = cfg->class_size[cfg->nclasses - 1];
	for (uint32_t c = 0; c < cfg->nclasses; c++) {
		uint32_t s = cfg->class_size[c];
		nh->class_size[c] = s;
		nh->class_cap[c] = sbsz / s;
		nh->class_magic[c] = ((1ull << 40) + s - 1) / s;
	}
	for (uint32_t c = cfg->nclasses; c < 32; c++)
		nh->class_size[c] = 0xffffffffu;
	for (uint32_t e = 0; e < 64; e++) {
		/* first class that can hold 2^(e-1)+1 bytes */
		uint64_t lo = e == 0 ? 1 : (1ull << (e - 1)) + 1;
		uint32_t c = 0;
		while (c < cfg->nclasses && cfg->class_size[c] < lo)
			c++;
		if (c >= cfg->nclasses)
			c = cfg->nclasses - 1;
		nh->class_for_log2[e] = (uint8_t)c;
	}
	uint64_t off = align_up(sizeof(nat_hdr), 256);
	nh->off_stats = off;
	off += NAT_STATS_BYTES;
	nh->off_state = off;
	off = align_up(off + nh->nsb * 8, 256);
	nh->off_next = off;
	off = align_up(off + nh->nsb * 4, 256);
	nh->off_bitmap = off;
	off = align_up(off + nh->nsb * nh->bm_words * 4ull, 256);
	nh->off_heads = off;
	off = align_up(off + (uint64_t)nh->nctx * NAT_HEADS_PER_CTX * 8, 256);
	nh->off_pool = off;
	off = align_up(off + (uint64_t)nh->npool * NAT_POOL_STRIDE_BYTES, 256);
	nh->off_pbitmap = off;
	off = align_up(off + ((nh->npages + 63) / 64) * 8, 256);
	nh->off_plen = off;
	off = align_up(off + nh->npages, 256);
	nh->off_psum = off;
	{
		/* summary words per level, then their top-level words per level */
		uint64_t ns = (((nh->npages + 63) / 64) + 63) / 64, ns2 = (ns + 63) / 64;
		off = align_up(off + NAT_PAGE_LEVELS * (ns + ns2) * 8, 256);
	}
	/* context spare lines: up to 16 empty superblocks per context, at most
	 * a quarter of the superblocks over all contexts */
	nh->off_spare = off;
	off = align_up(off + (uint64_t)nh->nctx * 128, 256);
	{
		uint64_t q = nh->nsb / (4ull * nh->nctx);
		nh->spare_slots = (uint32_t)(q < 16 ? q : 16);
	}
	/* partial index (detached, partly free superblocks waiting to be
	 * attached again): per XCD and class, a bit per superblock and two
	 * summary levels, each block on its own 128-B lines */
	{
		uint64_t w0 = (nh->nsb + 63) / 64, w1 = (w0 + 63) / 64, w2 = (w1 + 63) / 64;
		nh->pidx_words = (uint32_t)align_up(w0 + w1 + w2, 16);
	}
	nh->off_pidx = off;
	off = align_up(off + 8ull * nh->nclasses * nh->pidx_words * 8ull, 256);
	/* metadata bytes; data placed after allocation so it is absolutely
	 * aligned to the superblock and page sizes */
	nh->off_data = off;
	return 0;
}


uint64_t kma_native_alloc_bytes(const nat_hdr *nh)
{
	uint64_t sbsz = 1ull << nh->sb_log2, psz = 1ull << nh->page_log2;
	return nh->off_data + sbsz + (nh->nsb << nh->sb_log2) + psz + (nh->npages << nh->page_log2);
}

void kma_native_place(nat_hdr *nh, uint64_t base, uint64_t total)
{
	uint64_t sbsz = 1ull << nh->sb_log2, psz = 1ull << nh->page_log2;
	uint64_t meta = nh->off_data;
	nh->off_data = align_up(base + meta, sbsz) - base;
	nh->off_pages = align_up(base + nh->off_data + (nh->nsb << nh->sb_log2), psz) - base;
	nh->bytes = total;
}
