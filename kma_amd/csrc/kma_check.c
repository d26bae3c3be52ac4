/*
 * kma_check.c - heap walker / orphan detector (the reference's
 * commented-out heap walker, kma.cl:528-548, and its heap dumps,
 * tb_kma.c:304-313) for both layouts.  Works on an image taken while no
 * kernel uses the heap.  Native layout (kma_snapshot of bytes
 * [0, off_data)): every superblock must be in exactly one place (a pool
 * shard, a context spare line, a context head, the partial index of its
 * home XCD and class, or in use with live blocks), its count must equal its
 * live bits, and the page bitmap must match the run lengths and their
 * window summaries.  Reference layout (the
 * whole image): the free queue is walked as the reference's validator does
 * (tb_clIndexedQueue.c:131-168) and every superblock's state word, block
 * size and bitmap must agree with where it is (queued, a class head, or
 * detached with live blocks).
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
#define ST_HOME_SHIFT 38
#define ST_LISTED (1ull << 41)
#define ST_GEN_SHIFT 42
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

/* ---------------- reference layout ---------------- */
/* block size of class c and slots per superblock (kma.cl:55-114; the same
 * arithmetic as the device code in kma_ref_device.h) */
static uint32_t ref_size_by_sbid(uint32_t log2, int c)
{
	if (c < 0 || c > (int)log2 - 1)
		return 0;
	int e = c + 2;
	if (e <= (int)(log2 >> 1))
		return 1u << e;
	uint32_t items = (1u << log2) >> e;
	uint32_t ab = (items >> 3) + ((items & 7u) ? 1u : 0u);
	ab = (ab + 3u) & ~3u;
	return (((1u << log2) - 12u - ab) / items) & ~3u;
}

static uint32_t ref_slots_by_size(uint32_t log2, uint32_t size)
{
	uint32_t space = (1u << log2) - 12u;
	uint32_t s = space / size;
	if (s & 0x1fu)
		s += 32u;
	s = (s >> 3) & ~3u;
	return (space - s) / size;
}

/* 1 if the header is a reference-layout heap's (kma.h:81-85): `bytes`
 * = S superblocks + header, free.stride = log2 of the superblock size,
 * free.base != 0. */
static int ref_header(const uint8_t *img, uint64_t bytes, uint32_t *log2_out, uint64_t *S_out)
{
	if (bytes < 32)
		return 0;
	uint64_t hb = rd64(img, 0), log2 = rd64(img, 16);
	if (log2 < 8 || log2 > 16)
		return 0;
	uint64_t hdr = 32 + 8 * (log2 - 1);
	if (hb < hdr || hb > bytes || rd64(img, 8) == 0)
		return 0;
	uint64_t S = (hb - hdr) >> log2;
	if (S < 2 || S > 65536 || hb != (S << log2) + hdr)
		return 0;
	*log2_out = (uint32_t)log2;
	*S_out = S;
	return 1;
}

static int ref_image_check(const uint8_t *img, uint32_t log2, uint64_t S, kma_heap_report *out)
{
	rep p = {out, 0};
	const uint32_t buckets = log2 - 1u;
	const uint64_t hdr = 32 + 8ull * buckets, sbsz = 1ull << log2, sb0 = rd64(img, 8);
	uint8_t *inq = (uint8_t *)calloc(S, 1);
	uint8_t *head_of = (uint8_t *)calloc(S, 1); /* class + 1 */
	if (!inq || !head_of) {
		free(inq);
		free(head_of);
		set_err("host allocation failed");
		return -1;
	}
	/* the free queue, from its head: nodes are 1-based indexes in bits
	 * 1..16 of the links (clIndexedQueue.cl:89-91) */
	uint32_t headw = rd32(img, 24), tailw = rd32(img, 28);
	uint32_t hi = (headw & 0x1fffeu) >> 1, ti = (tailw & 0x1fffeu) >> 1;
	if ((headw | tailw) & 1u)
		problem(&p, "queue head %#x / tail %#x carries the poison bit", headw, tailw);
	if (ti == 0 || ti > S)
		problem(&p, "queue tail names node %u of %llu", ti, (unsigned long long)S);
	if (hi == 0 || hi > S) {
		problem(&p, "queue head names node %u of %llu", hi, (unsigned long long)S);
	} else {
		uint64_t n = 0;
		for (uint32_t i = hi;;) {
			if (inq[i - 1u]) {
				problem(&p, "queue links back to superblock %u (a cycle)", i - 1u);
				break;
			}
			inq[i - 1u] = 1;
			n++;
			uint32_t nw = rd32(img, hdr + (uint64_t)(i - 1u) * sbsz);
			if (nw & 1u)
				problem(&p, "queued superblock %u has a poisoned link %#x", i - 1u, nw);
			uint32_t ni = (nw & 0x1fffeu) >> 1;
			if (ni == 0) {
				if (i != ti)
					problem(&p, "queue ends at superblock %u but its tail names %u", i - 1u, ti - 1u);
				break;
			}
			if (ni > S) {
				problem(&p, "queued superblock %u links to node %u, past the heap", i - 1u, ni);
				break;
			}
			i = ni;
		}
		out->sb_free = n;
	}
	/* class heads */
	for (uint32_t c = 0; c < buckets; c++) {
		uint64_t v = rd64(img, 32 + 8ull * c);
		if (v == 0)
			continue;
		if (v == 1) {
			problem(&p, "class %u head is POISON (a refill did not finish)", c);
			continue;
		}
		if (v < sb0 || (v - sb0) % sbsz || (v - sb0) / sbsz >= S) {
			problem(&p, "class %u head %#llx is not a superblock", c, (unsigned long long)v);
			continue;
		}
		uint64_t s = (v - sb0) / sbsz;
		if (head_of[s])
			problem(&p, "superblock %llu heads classes %u and %u", (unsigned long long)s, head_of[s] - 1u, c);
		head_of[s] = (uint8_t)(c + 1u);
		out->sb_heads++;
		uint32_t size = rd32(img, hdr + s * sbsz + 8);
		if (size != ref_size_by_sbid(log2, (int)c))
			problem(&p, "class %u head: superblock %llu holds %u-byte blocks", c, (unsigned long long)s, size);
	}
	/* superblocks */
	for (uint64_t s = 0; s < S; s++) {
		const uint64_t so = hdr + s * sbsz;
		uint32_t state = rd32(img, so + 4), size = rd32(img, so + 8);
		uint32_t slots = state >> 16, fr = state & 0xffffu;
		uint64_t live = 0, beyond = 0;
		uint32_t words = (slots + 31u) / 32u;
		if (words * 4ull > sbsz - 12)
			words = 0; /* a corrupt slot count: reported below */
		for (uint32_t k = 0; k < words; k++) {
			uint32_t b = rd32(img, so + sbsz - 4ull * (k + 1u));
			live += (uint64_t)__builtin_popcount(b);
			for (uint32_t j = 0; j < 32; j++)
				if (((b >> j) & 1u) && k * 32u + j >= slots)
					beyond++;
		}
		if (inq[s]) {
			if (head_of[s])
				problem(&p, "queued superblock %llu is also the head of class %u", (unsigned long long)s,
					head_of[s] - 1u);
			if (fr || live)
				problem(&p, "queued superblock %llu: free count %u, %llu bits set", (unsigned long long)s, fr,
					(unsigned long long)live);
			continue;
		}
		if (s == S - 1 && state == 0 && size == 0 && !head_of[s] && rd32(img, so) == 0) {
			out->sb_unqueued++; /* never queued by the init (kma.cl:42-51) */
			continue;
		}
		int c = -1;
		for (uint32_t k = 0; k < buckets && c < 0; k++)
			if (size && ref_size_by_sbid(log2, (int)k) == size)
				c = (int)k;
		if (c < 0) {
			problem(&p, "superblock %llu is in no queue and holds %u-byte blocks, no class's size",
				(unsigned long long)s, size);
			continue;
		}
		uint32_t want = ref_slots_by_size(log2, size);
		if (slots != want) {
			problem(&p, "superblock %llu of class %d: %u slots, the class has %u", (unsigned long long)s, c,
				slots, want);
			continue;
		}
		if (beyond)
			problem(&p, "superblock %llu has %llu bits past its %u slots", (unsigned long long)s,
				(unsigned long long)beyond, slots);
		if ((uint64_t)fr + live != slots)
			problem(&p, "superblock %llu counts %u free slots, %llu of %u bits are set", (unsigned long long)s, fr,
				(unsigned long long)live, slots);
		if (!head_of[s]) {
			if (live == 0) {
				out->orphans++;
				problem(&p, "superblock %llu of class %d has no live block, is no class head and is not "
					    "queued (orphan)",
					(unsigned long long)s, c);
			} else {
				out->sb_detached++;
			}
		}
		out->live_blocks += live;
		out->live_bytes += live * size;
	}
	out->sb_total = S;
	free(inq);
	free(head_of);
	return 0;
}

int kma_image_check(const void *image, uint64_t bytes, kma_heap_report *out)
{
	if (!image || !out || bytes < 32) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	const uint8_t *img = (const uint8_t *)image;
	nat_hdr h;
	if (bytes < sizeof(nat_hdr) || (memcpy(&h, img, sizeof(h)), h.magic != NAT_MAGIC)) {
		uint32_t log2;
		uint64_t S;
		if (ref_header(img, bytes, &log2, &S))
			return ref_image_check(img, log2, S, out);
		set_err("neither a native nor a reference-layout heap image (or truncated)");
		return -1;
	}
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
		{h.off_pidx, 8ull * h.nclasses * h.pidx_words * 8},
	};
	int bad = h.nsb > 0xfffffff0ull || h.npages > (1ull << 40) || h.bm_words > 4096 || h.npool > 512 ||
		  h.nctx > 65536;
	const uint64_t pw0 = (h.nsb + 63) >> 6, pw1 = (pw0 + 63) >> 6, pw2 = (pw1 + 63) >> 6;
	bad = bad || h.pidx_words < pw0 + pw1 + pw2;
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
	uint8_t *listed = (uint8_t *)calloc(nsb + 1, 1);    /* partial-index bits naming it */
	uint16_t *lplace = (uint16_t *)calloc(nsb + 1, 2); /* (XCD << 8) | class of that index */
	if (!where || !headed || !listed || !lplace) {
		free(where);
		free(headed);
		free(listed);
		free(lplace);
		set_err("host allocation failed");
		return -1;
	}
	/* partial indexes: level-0 bits name listed superblocks; a non-zero
	 * word must have its bit set one level up (set bits are hints) */
	for (uint32_t x = 0; x < 8; x++)
		for (uint32_t c = 0; c < h.nclasses; c++) {
			const uint64_t P = h.off_pidx + ((uint64_t)x * h.nclasses + c) * h.pidx_words * 8;
			for (uint64_t i = 0; i < pw0; i++) {
				uint64_t w = rd64(img, P + i * 8);
				if (!w)
					continue;
				if (!((rd64(img, P + (pw0 + (i >> 6)) * 8) >> (i & 63)) & 1ull))
					problem(&p, "partial index of XCD %u class %u: word %llu is not zero but its summary bit is "
						    "clear",
						x, c, (unsigned long long)i);
				for (; w; w &= w - 1) {
					uint64_t sb = i * 64 + (uint64_t)__builtin_ctzll(w);
					if (sb >= nsb) {
						problem(&p, "partial index of XCD %u class %u names superblock %llu, past the heap", x,
							c, (unsigned long long)sb);
						continue;
					}
					if (listed[sb] < 255)
						listed[sb]++;
					lplace[sb] = (uint16_t)((x << 8) | c);
				}
			}
			for (uint64_t i = 0; i < pw1; i++)
				if (rd64(img, P + (pw0 + i) * 8) && !((rd64(img, P + (pw0 + pw1 + (i >> 6)) * 8) >> (i & 63)) & 1ull))
					problem(&p, "partial index of XCD %u class %u: summary word %llu is not zero but its top bit "
						    "is clear",
						x, c, (unsigned long long)i);
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
		if (listed[sb] > 1 || (listed[sb] != 0) != ((st & ST_LISTED) != 0))
			problem(&p, "superblock %llu: listed bit %d, in %u partial indexes", (unsigned long long)sb,
				(st & ST_LISTED) != 0, listed[sb]);
		if (c1 == 0) {
			if (cnt || live || (st & (ST_OWNED | ST_LISTED)))
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
		if (st & ST_LISTED) {
			const uint32_t home = (uint32_t)(st >> ST_HOME_SHIFT) & 7u;
			if (st & ST_OWNED)
				problem(&p, "superblock %llu is listed and owned", (unsigned long long)sb);
			else if (listed[sb] == 1 && lplace[sb] != (uint16_t)((home << 8) | (c1 - 1u)))
				problem(&p, "superblock %llu of XCD %u class %u is in the partial index of XCD %u class %u",
					(unsigned long long)sb, home, c1 - 1u, lplace[sb] >> 8, lplace[sb] & 0xffu);
			out->sb_listed++;
		} else if (!(st & ST_OWNED)) {
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
	free(listed);
	free(lplace);
	return 0;
}

int kma_heap_check(kma_heap *h, kma_heap_report *out)
{
	if (!h || !out) {
		set_err("bad arguments");
		return -1;
	}
	if (h->cfg.layout == KMA_LAYOUT_REFERENCE) {
		uint8_t *img = (uint8_t *)malloc(h->bytes);
		uint64_t tw[REF_T_WORDS];
		if (!img) {
			set_err("host allocation failed");
			return -1;
		}
		int rc = -1;
		if (kma_snapshot(h, 0, img, h->bytes) == 0 && kma_i_read_ref_trailer(h, tw) == 0) {
			rc = kma_image_check(img, h->bytes, out);
			for (uint32_t k = 0; rc == 0 && k < REF_T_WORDS; k++)
				out->counters[k] = tw[k];
		}
		free(img);
		return rc;
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
