/*
 * kma_check.c - heap walker / orphan detector for the native layout (the
 * reference's commented-out heap walker, kma.cl:528-548, and its heap
 * dumps, tb_kma.c:304-313).  Works on a metadata image (kma_snapshot of
 * bytes [0, off_data)) taken while no kernel uses the heap: every
 * superblock must be in exactly one place (a pool shard, a context spare
 * line, a context head, or in use with live blocks), its count must equal
 * its live bits, and the page bitmap must match the run lengths and their
 * window summaries.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_internal.h"
#include "kma_work.h"

/* superblock state word (kma_device.h) */
#define ST_COUNT 0xffffffffull
#define ST_CLASS_SHIFT 32
#define ST_OWNED (1ull << 37)
#define ST_GEN_SHIFT 41
#define POOL_STRIDE_WORDS 16u
#define SPARE_WORDS 32u /* u32 per context spare line */

typedef struct {
	kma_heap_report *r;
	size_t used;
} rep;

static void problem(rep *p, const char *fmt, ...)
{
	p->r->problems++;
	if (p->used + 2 >= sizeof(p->r->first))
		return;
	va_list ap;
	va_start(ap, fmt);
	int n = vsnprintf(p->r->first + p->used, sizeof(p->r->first) - p->used - 1, fmt, ap);
	va_end(ap);
	if (n < 0)
		return;
	p->used += (size_t)n < sizeof(p->r->first) - p->used - 1 ? (size_t)n : sizeof(p->r->first) - p->used - 1;
	p->r->first[p->used++] = '\n';
	p->r->first[p->used] = 0;
}

static uint64_t rd64(const uint8_t *img, uint64_t off)
{
	uint64_t v;
	memcpy(&v, img + off, 8);
	return v;
}

static uint32_t rd32(const uint8_t *img, uint64_t off)
{
	uint32_t v;
	memcpy(&v, img + off, 4);
	return v;
}

static uint64_t page_valid(uint64_t npages, uint64_t wi)
{
	uint64_t n = npages - (wi << 6);
	return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

/* bit j: pages j .. j + 2^l - 1 free and valid, j a multiple of 2^l */
static uint64_t page_windows(uint64_t w, uint32_t l, uint64_t valid)
{
	uint64_t z = ~w & valid;
	for (uint32_t sh = 1; sh < (1u << l); sh <<= 1)
		z &= z >> sh;
	uint64_t pat = 1;
	for (uint32_t sh = 1u << l; sh < 64u; sh <<= 1)
		pat |= pat << sh;
	return z & pat;
}

int kma_image_check(const void *image, uint64_t bytes, kma_heap_report *out)
{
	if (!image || !out || bytes < sizeof(nat_hdr)) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	const uint8_t *img = (const uint8_t *)image;
	nat_hdr h;
	memcpy(&h, img, sizeof(h));
	if (h.magic != NAT_MAGIC || h.off_data > bytes || h.nctx == 0 || h.npool < 8 || h.nclasses == 0 ||
	    h.nclasses > 32 || h.spare_slots > 16) {
		set_err("not a native heap image (or truncated before its data offset)");
		return -1;
	}
	/* every table must lie inside the metadata (a corrupt header must not
	 * send the walk outside the image) */
	const uint64_t nw = (h.npages + 63) >> 6, ns_ = (nw + 63) >> 6, ns2_ = (ns_ + 63) >> 6;
	const struct {
		uint64_t off, len;
	} tab[] = {
		{h.off_state, h.nsb * 8},
		{h.off_next, h.nsb * 4},
		{h.off_bitmap, h.nsb * h.bm_words * 4},
		{h.off_heads, (uint64_t)h.nctx * NAT_HEADS_PER_CTX * 8},
		{h.off_pool, (uint64_t)h.npool * NAT_POOL_STRIDE_BYTES},
		{h.off_pbitmap, nw * 8},
		{h.off_plen, h.npages},
		{h.off_stats, NAT_STATS_BYTES},
		{h.off_psum, (uint64_t)NAT_PAGE_LEVELS * (ns_ + ns2_) * 8},
		{h.off_spare, (uint64_t)h.nctx * 128},
	};
	int bad = h.nsb > 0xfffffff0ull || h.npages > (1ull << 40) || h.bm_words > 4096 || h.npool > 512 ||
		  h.nctx > 65536;
	for (unsigned i = 0; i < sizeof(tab) / sizeof(tab[0]) && !bad; i++)
		bad = tab[i].off > h.off_data || tab[i].len > h.off_data - tab[i].off;
	if (bad) {
		set_err("native heap image with tables outside its metadata");
		return -1;
	}
	rep p = {out, 0};
	const uint64_t nsb = h.nsb, W = h.bm_words, S = h.npool / 8;
	uint8_t *where = (uint8_t *)calloc(nsb + 1, 1); /* free places holding each superblock */
	uint8_t *headed = (uint8_t *)calloc(nsb + 1, 1);
	if (!where || !headed) {
		free(where);
		free(headed);
		set_err("host allocation failed");
		return -1;
	}
	/* pool shards */
	for (uint32_t s = 0; s < h.npool; s++) {
		uint64_t hp = h.off_pool + (uint64_t)s * POOL_STRIDE_WORDS * 8;
		uint32_t top = (uint32_t)rd64(img, hp);
		uint64_t count = rd64(img, hp + 8), walked = 0;
		while (top && walked <= nsb) {
			if (top - 1u >= nsb) {
				problem(&p, "shard %u links superblock %u, past the heap", s, top - 1u);
				break;
			}
			if (where[top - 1u] < 255)
				where[top - 1u]++;
			walked++;
			top = rd32(img, h.off_next + (uint64_t)(top - 1u) * 4);
		}
		if (walked > nsb)
			problem(&p, "shard %u: its chain does not end", s);
		if (walked != count)
			problem(&p, "shard %u holds %llu superblocks, its count says %llu", s, (unsigned long long)walked,
				(unsigned long long)count);
		uint64_t mask = rd64(img, h.off_stats + 256 + 128ull * (s / S));
		if (walked && !((mask >> (s % S)) & 1ull))
			problem(&p, "shard %u is not empty but its mask bit is clear", s);
		out->sb_free += walked;
	}
	/* context spare lines */
	for (uint32_t c = 0; c < h.nctx; c++)
		for (uint32_t k = 0; k < SPARE_WORDS; k++) {
			uint32_t v = rd32(img, h.off_spare + (uint64_t)c * 128 + 4ull * k);
			if (!v)
				continue;
			if (k >= h.spare_slots || v - 1u >= nsb) {
				problem(&p, "context %u spare word %u holds %u", c, k, v);
				continue;
			}
			if (where[v - 1u] < 255)
				where[v - 1u]++;
			out->sb_in_spares++;
			out->sb_free++;
		}
	/* context heads */
	for (uint32_t c = 0; c < h.nctx; c++)
		for (uint32_t k = 0; k < NAT_HEADS_PER_CTX; k++) {
			uint64_t hv = rd64(img, h.off_heads + ((uint64_t)c * NAT_HEADS_PER_CTX + k) * 8);
			if (hv == 0)
				continue;
			uint32_t sb = (uint32_t)hv - 2u;
			if (k >= h.nclasses || hv == 1 || sb >= nsb) {
				problem(&p, "context %u head %u = %#llx", c, k, (unsigned long long)hv);
				continue;
			}
			uint64_t st = rd64(img, h.off_state + (uint64_t)sb * 8);
			if (((st >> ST_CLASS_SHIFT) & 0x1f) != k + 1u || !(st & ST_OWNED) ||
			    (uint32_t)(st >> ST_GEN_SHIFT) != (uint32_t)(hv >> 32))
				problem(&p, "context %u head %u names superblock %u in state %#llx", c, k, sb,
					(unsigned long long)st);
			if (headed[sb] < 255)
				headed[sb]++;
			out->sb_heads++;
		}
	/* superblocks */
	for (uint64_t sb = 0; sb < nsb; sb++) {
		uint64_t st = rd64(img, h.off_state + sb * 8);
		uint32_t cnt = (uint32_t)(st & ST_COUNT), c1 = (uint32_t)((st >> ST_CLASS_SHIFT) & 0x1f);
		uint64_t live = 0, beyond = 0;
		uint32_t cap = c1 && c1 <= h.nclasses ? h.class_cap[c1 - 1u] : 0u;
		for (uint64_t w = 0; w < W; w++) {
			uint32_t b = rd32(img, h.off_bitmap + (sb * W + w) * 4);
			live += (uint64_t)__builtin_popcount(b);
			for (uint32_t j = 0; j < 32; j++)
				if (((b >> j) & 1u) && w * 32 + j >= cap)
					beyond++;
		}
		if (c1 == 0) {
			if (cnt || live || (st & ST_OWNED))
				problem(&p, "free superblock %llu: state %#llx, %llu bits set", (unsigned long long)sb,
					(unsigned long long)st, (unsigned long long)live);
			if (where[sb] != 1)
				problem(&p, "free superblock %llu is in %u free lists (orphan if 0)", (unsigned long long)sb,
					where[sb]);
			if (headed[sb])
				problem(&p, "free superblock %llu is a context head", (unsigned long long)sb);
			continue;
		}
		if (c1 > h.nclasses)
			problem(&p, "superblock %llu has class %u", (unsigned long long)sb, c1 - 1u);
		if (where[sb])
			problem(&p, "superblock %llu of class %u is also in a free list", (unsigned long long)sb, c1 - 1u);
		if (beyond)
			problem(&p, "superblock %llu has %llu bits past its %u slots", (unsigned long long)sb,
				(unsigned long long)beyond, cap);
		if (cnt != live)
			problem(&p, "superblock %llu counts %u blocks, %llu bits are set", (unsigned long long)sb, cnt,
				(unsigned long long)live);
		if (((st & ST_OWNED) != 0) != (headed[sb] == 1))
			problem(&p, "superblock %llu: owned bit %d, named by %u heads", (unsigned long long)sb,
				(st & ST_OWNED) != 0, headed[sb]);
		if (!(st & ST_OWNED)) {
			if (cnt == 0) {
				out->orphans++;
				problem(&p, "superblock %llu of class %u has no live block, no owner and is in no free "
					    "list (orphan)",
					(unsigned long long)sb, c1 - 1u);
			} else {
				out->sb_detached++;
			}
		}
		out->live_blocks += live;
		out->live_bytes += live * (c1 <= h.nclasses ? h.class_size[c1 - 1u] : 0u);
	}
	/* pages: the bitmap must be a sequence of runs, each starting at a page
	 * whose plen gives its length (plen of freed runs is left stale, so only
	 * entries at set bits are read), and the window summaries must be exact */
	const uint64_t nwords = (h.npages + 63) >> 6, ns = (nwords + 63) >> 6, ns2 = (ns + 63) >> 6;
	for (uint64_t wi = 0; wi < ns * 64; wi++) {
		uint64_t w = wi < nwords ? rd64(img, h.off_pbitmap + wi * 8) : 0;
		for (uint32_t j = 0; j < 64 && wi < nwords;) {
			if (!((w >> j) & 1ull)) {
				j++;
				continue;
			}
			uint32_t k = img[h.off_plen + (wi << 6) + j];
			uint64_t run = k >= 64 ? ~0ull : ((1ull << k) - 1ull);
			if (k == 0 || j + k > 64 || ((run << j) & ~w) != 0) {
				problem(&p, "page %llu is in use but no run of %u pages starts there",
					(unsigned long long)((wi << 6) + j), k);
				break;
			}
			out->page_runs++;
			out->pages_used += k;
			j += k;
		}
		if (wi < nwords && (w & ~page_valid(h.npages, wi)))
			problem(&p, "page word %llu has bits past the last page", (unsigned long long)wi);
		for (uint32_t l = 0; l < NAT_PAGE_LEVELS; l++) {
			uint64_t sw = rd64(img, h.off_psum + ((uint64_t)l * ns + (wi >> 6)) * 8);
			int full = wi >= nwords || page_windows(w, l, page_valid(h.npages, wi)) == 0;
			if (((sw >> (wi & 63)) & 1ull) != (uint64_t)full)
				problem(&p, "page word %llu level %u: summary bit %d, free windows %s", (unsigned long long)wi,
					l, (int)((sw >> (wi & 63)) & 1ull), full ? "none" : "some");
		}
	}
	for (uint32_t l = 0; l < NAT_PAGE_LEVELS && ns; l++)
		for (uint64_t si = 0; si < ns2 * 64; si++) {
			uint64_t top = rd64(img, h.off_psum + ((uint64_t)NAT_PAGE_LEVELS * ns + (uint64_t)l * ns2 + (si >> 6)) * 8);
			int full = si >= ns || rd64(img, h.off_psum + ((uint64_t)l * ns + si) * 8) == ~0ull;
			if (((top >> (si & 63)) & 1ull) != (uint64_t)full)
				problem(&p, "summary word %llu level %u: top-level bit %d, word %s", (unsigned long long)si, l,
					(int)((top >> (si & 63)) & 1ull), full ? "full" : "not full");
		}
	out->pages_total = h.npages;
	out->sb_total = nsb;
	for (uint32_t k = 0; k < NAT_NSTATS && k < sizeof(out->counters) / 8; k++)
		out->counters[k] = rd64(img, h.off_stats + 8ull * k);
	free(where);
	free(headed);
	return 0;
}

int kma_heap_check(kma_heap *h, kma_heap_report *out)
{
	if (!h || !out || h->cfg.layout != KMA_LAYOUT_NATIVE) {
		set_err("bad arguments (a native heap)");
		return -1;
	}
	uint64_t n = h->nh.off_data;
	uint8_t *img = (uint8_t *)malloc(n);
	if (!img) {
		set_err("host allocation failed");
		return -1;
	}
	if (kma_snapshot(h, 0, img, n) != 0) {
		free(img);
		return -1;
	}
	int rc = kma_image_check(img, n, out);
	free(img);
	return rc;
}
