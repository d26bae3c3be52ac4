/*
 * refmodel.c - CPU restatement of RSpliet/KMA's superblock allocator.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, the smoke check and the
 * bench's CPU-baseline leg load this; libkma.so never does.
 *
 * Byte layout reproduced (64-bit device, kma.h:71-85, clIndexedQueue.h:38-51):
 *   struct clheap  @0:  u64 bytes | u64 free.base | u64 free.stride |
 *                       u32 free.head | u32 free.tail | u64 sb[BUCKETS]
 *   superblock i   @hdr + i*SB:  u32 q.next | u32 state | u32 size | data...
 *   allocation bitmap: bit b lives in the u32 at SB - 4*((b>>5)+1), i.e. the
 *   bitmap grows downwards from the end of the superblock (kma.cl:184-188,
 *   244-250, 336-337).
 * Queue words (clIndexedQueue.cl:39-42,89-91): bit 0 poison, bits 1..16 index
 * (index = ((addr-base)>>stride)+1), bits 22..32 an 11-bit tag.
 */
#include "refmodel.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define POISON 1u

/* ---------------- little-endian image accessors ---------------- */
static inline uint32_t rd32(const rm_heap *h, uint64_t addr)
{
	uint32_t v;
	memcpy(&v, h->img + (addr - h->base), 4);
	return v;
}
static inline void wr32(rm_heap *h, uint64_t addr, uint32_t v)
{
	memcpy(h->img + (addr - h->base), &v, 4);
}
static inline uint64_t rd64(const rm_heap *h, uint64_t addr)
{
	uint64_t v;
	memcpy(&v, h->img + (addr - h->base), 8);
	return v;
}
static inline void wr64(rm_heap *h, uint64_t addr, uint64_t v)
{
	memcpy(h->img + (addr - h->base), &v, 8);
}

/* Header field addresses (kma.h:81-85 with clIndexedQueue at offset 8). */
#define H_BYTES(h)   ((h)->base + 0)
#define Q_BASE(h)    ((h)->base + 8)
#define Q_STRIDE(h)  ((h)->base + 16)
#define Q_HEAD(h)    ((h)->base + 24)
#define Q_TAIL(h)    ((h)->base + 28)
#define H_SB(h, c)   ((h)->base + 32 + 8 * (uint64_t)(c))

/* Tagged index word (clIndexedQueue.cl:89-91). */
static inline uint32_t q_ptr(uint32_t i, uint32_t t)
{
	return (i & 0x1fffeu) | (((t + 1u) & 0x7ffu) << 22);
}
static inline uint32_t q_tag(uint32_t i) { return (i >> 22) & 0x7ffu; }
static inline uint32_t q_idx(uint32_t i) { return (i & 0x1fffeu) >> 1; }

/* ---------------- size classes (kma.cl:55-148) ---------------- */
uint32_t rm_header_bytes(uint32_t log2)
{
	return 32u + 8u * (log2 - 1u);
}

uint32_t rm_size_by_sbid(uint32_t log2, int sbid)
{
	/* kma.cl:55-90: small classes are plain powers of two; from
	 * 2^(log2/2) upwards the block size is stretched to use the space
	 * left after the 12-byte header and the rounded bitmap. */
	uint32_t sbsize = 1u << log2;
	int buckets = (int)log2 - 1;
	if (sbid > buckets || sbid < 0)
		return 0;
	int e = sbid + 2;
	if (e <= (int)(log2 >> 1))
		return 1u << e;
	uint32_t bytes = 1u << e;
	uint32_t items = sbsize / bytes;
	uint32_t ab = (items >> 3) + ((items & 7u) ? 1u : 0u);
	ab += (ab & 3u) ? (4u - (ab & 3u)) : 0u;
	uint32_t total = sbsize - (12u + ab);
	return (total / items) & ~3u;
}

int rm_slots_by_size(uint32_t log2, uint32_t size)
{
	/* kma.cl:96-114: first estimate, reserve whole 32-bit bitmap words
	 * for it, recompute. */
	uint32_t space = (1u << log2) - 12u;
	int slots = (int)(space / size);
	if (slots & 0x1f)
		slots += 32;
	slots >>= 3;
	slots &= ~0x3;
	space -= (uint32_t)slots;
	return (int)(space / size);
}

int rm_sbid_by_size(uint32_t log2, uint64_t size)
{
	/* kma.cl:119-148 */
	int buckets = (int)log2 - 1;
	if (size > (1ull << log2) - 16u)
		return -1;
	if (size == 0)
		return -1;
	if (size < 4)
		return 0;
	uint64_t ssize = size >> 2;
	int sbid = 0;
	for (int i = 0; i < buckets; i++, ssize >>= 1)
		if (ssize & 1)
			sbid = i;
	for (;;) {
		int test = (int)rm_size_by_sbid(log2, sbid);
		if ((int64_t)size > test && test > 0)
			sbid++;
		else if ((int64_t)size < (int64_t)rm_size_by_sbid(log2, sbid - 1))
			sbid--;
		else
			return sbid;
	}
}

/* ---------------- indexed queue, serialized (clIndexedQueue.cl) ---------- */
static uint32_t q_ptr2idx(const rm_heap *h, uint64_t addr)
{
	uint64_t qb = rd64(h, Q_BASE(h));
	uint32_t stride = (uint32_t)rd64(h, Q_STRIDE(h));
	if (addr < qb)
		return 0;
	uint64_t idx = addr - qb;
	if (idx & ((1ull << stride) - 1))
		return 0;
	idx >>= stride;
	idx++;
	if (idx > ((1u << 20) - 1))
		return 0;
	return (uint32_t)idx << 1;
}

static uint64_t q_idx2ptr(const rm_heap *h, uint32_t v)
{
	uint64_t qb = rd64(h, Q_BASE(h));
	uint32_t stride = (uint32_t)rd64(h, Q_STRIDE(h));
	uint32_t i = (v >> 1) & 0xfffffu;
	if (i == 0)
		return 0;
	return qb + ((uint64_t)(i - 1) << stride);
}

static void q_enqueue(rm_heap *h, uint64_t item)
{
	/* clIndexedQueue.cl:119-166 with no concurrent actor: the tail is
	 * never lagging, so the link CAS succeeds on the first pass. */
	uint32_t tag = q_tag(rd32(h, item));
	wr32(h, item, q_ptr(0, tag - 1u));
	uint32_t idx = q_ptr2idx(h, item);
	uint32_t tailidx = rd32(h, Q_TAIL(h));
	uint64_t tail = q_idx2ptr(h, tailidx);
	uint32_t nextidx = rd32(h, tail);
	if (q_idx(nextidx) == 0) {
		wr32(h, tail, q_ptr(idx, q_tag(nextidx)));
	} else {
		/* Only reachable with a lagging tail; kept for completeness. */
		wr32(h, Q_TAIL(h), q_ptr(nextidx, q_tag(tailidx)));
		q_enqueue(h, item);
		return;
	}
	wr32(h, Q_TAIL(h), q_ptr(idx, q_tag(tailidx)));
}

static uint64_t q_dequeue(rm_heap *h)
{
	/* clIndexedQueue.cl:173-209: returns the old head node (the dummy);
	 * its successor becomes the new dummy. */
	for (;;) {
		uint32_t headidx = rd32(h, Q_HEAD(h));
		uint64_t head = q_idx2ptr(h, headidx);
		uint32_t tailidx = rd32(h, Q_TAIL(h));
		uint32_t nextidx = rd32(h, head);
		if (q_idx(headidx) == q_idx(tailidx)) {
			if (q_idx(nextidx) == 0)
				return 0;
			wr32(h, Q_TAIL(h), q_ptr(nextidx, q_tag(tailidx)));
			continue;
		}
		wr32(h, Q_HEAD(h), q_ptr(nextidx, q_tag(headidx)));
		return head;
	}
}

int64_t rm_queue_walk(const rm_heap *h)
{
	uint32_t headv = rd32(h, Q_HEAD(h)), tailv = rd32(h, Q_TAIL(h));
	uint64_t node = q_idx2ptr(h, headv);
	uint64_t tail = q_idx2ptr(h, tailv);
	for (uint64_t count = 0; count <= (uint64_t)h->nsb + 1; count++) {
		uint32_t nx = rd32(h, node);
		if (q_idx(nx) == 0)
			return node == tail ? (int64_t)count + 1 : -2;
		node = q_idx2ptr(h, nx);
	}
	return -1;
}

/* ---------------- heap create + init ---------------- */
uint64_t rm_heap_bytes(uint32_t log2, uint32_t sblocks)
{
	/* kma.c:37 computes this in 32 bits; the model uses 64 bits so that
	 * 4 GiB-class heaps stay representable. */
	return ((uint64_t)sblocks << log2) + rm_header_bytes(log2);
}

int rm_init(rm_heap *h, uint32_t log2, uint32_t sblocks, uint64_t base,
	    uint8_t *img)
{
	memset(h, 0, sizeof(*h));
	if (sblocks < 2 || log2 < 6 || log2 > 20)
		return RM_EINVAL;
	/* The 16-bit effective queue index (clIndexedQueue.cl:89) caps the
	 * reference at 65,536 superblocks. */
	if (sblocks > 65536)
		return RM_EINVAL;
	h->log2 = log2;
	h->sb = 1u << log2;
	h->buckets = log2 - 1;
	h->hdr = rm_header_bytes(log2);
	h->nsb = sblocks;
	h->bytes = rm_heap_bytes(log2, sblocks);
	h->base = base;
	if (img) {
		h->img = img;
	} else {
		h->img = (uint8_t *)calloc(1, h->bytes);
		if (!h->img)
			return RM_EINVAL;
		h->owns_img = 1;
	}
	/* kma.c:37-44: the host writes only `bytes`. */
	wr64(h, H_BYTES(h), h->bytes);
	/* clheap_init, kma.cl:27-52 */
	for (uint32_t i = 0; i < h->buckets; i++)
		wr64(h, H_SB(h, i), 0);
	uint32_t pages = (uint32_t)(h->bytes >> log2) - 1u;
	uint64_t sb0 = h->base + h->hdr;
	/* clIndexedQueue_init(&heap->free, &sb[0], LOG2, &sb[0]),
	 * clIndexedQueue.cl:93-111 */
	wr64(h, Q_BASE(h), sb0);
	wr64(h, Q_STRIDE(h), log2);
	uint32_t tag = q_tag(rd32(h, sb0));
	wr32(h, sb0, q_ptr(0, tag));
	uint32_t idx = q_ptr2idx(h, sb0);
	wr32(h, Q_HEAD(h), q_ptr(idx, tag));
	wr32(h, Q_TAIL(h), q_ptr(idx, tag));
	for (uint32_t i = 1; i < pages; i++)
		q_enqueue(h, sb0 + ((uint64_t)i << log2));
	return RM_OK;
}

void rm_release(rm_heap *h)
{
	if (h->owns_img)
		free(h->img);
	h->img = NULL;
}

/* ---------------- malloc / free, serialized ---------------- */
uint64_t rm_malloc(rm_heap *h, uint64_t size, int *status)
{
	if (status)
		*status = RM_OK;
	int block = rm_sbid_by_size(h->log2, size);
	if (block < 0 || block >= (int)h->buckets)
		return 0;

	/* _kma_reserve_block, kma.cl:150-225 */
	uint64_t sb;
	uint32_t slot;
	uint64_t cursor = rd64(h, H_SB(h, block));
	if (cursor == 0) {
		sb = q_dequeue(h);
		if (!sb)
			return 0;
		uint32_t bsize = rm_size_by_sbid(h->log2, block);
		wr32(h, sb + 8, bsize);
		uint32_t slots = (uint32_t)rm_slots_by_size(h->log2, bsize);
		wr32(h, sb + 4, (slots << 16) | (slots - 1u));
		for (uint32_t i = 0; i < slots; i += 32)
			wr32(h, sb + h->sb - 4u * (i / 32u + 1u), 0);
		wr64(h, H_SB(h, block), sb);
		slot = slots - 1u;
	} else {
		uint32_t st = rd32(h, cursor + 4);
		uint32_t fr = st & 0xffffu;
		if (fr == 0) {
			/* A 1-slot class attached with no free slot:
			 * the reference spins here until a free. */
			if (status)
				*status = RM_HANG;
			return 0;
		}
		fr--;
		slot = fr;
		wr32(h, cursor + 4, (st & 0xffff0000u) | fr);
		if (fr == 0)
			wr64(h, H_SB(h, block), 0);
		sb = cursor;
	}

	/* _kma_get_block, kma.cl:232-273: first zero bit at or above `slot`,
	 * wrapping to 0; the ticket is written into the block. */
	uint32_t slots = rd32(h, sb + 4) >> 16;
	uint32_t bsize = rd32(h, sb + 8);
	uint32_t start = slot;
	for (int pass = 0; pass < 2; pass++) {
		for (uint32_t b = (pass == 0 ? start : 0); b < slots; b++) {
			uint64_t wa = sb + h->sb - 4u * ((b >> 5) + 1u);
			uint32_t w = rd32(h, wa);
			if (!((w >> (b & 31u)) & 1u)) {
				wr32(h, wa, w | (1u << (b & 31u)));
				uint64_t p = sb + 12u + (uint64_t)b * bsize;
				wr32(h, p, start);
				return p;
			}
		}
	}
	if (status)
		*status = RM_HANG;
	return 0;
}

void rm_free(rm_heap *h, uint64_t addr)
{
	/* kma.cl:305-373 */
	if (addr == 0)
		return;
	uint64_t first_sb = h->base + h->hdr;
	uint64_t off = addr - first_sb;
	uint64_t sb = first_sb + (off & ~(uint64_t)(h->sb - 1u));
	uint32_t size = rd32(h, sb + 8);
	uint64_t block = (addr - sb - 12u) / size;
	uint64_t wa = sb + h->sb - 4u * ((block >> 5) + 1u);
	uint32_t old = rd32(h, wa);
	wr32(h, wa, old & ~(1u << (block & 31u)));
	wr32(h, addr, old);
	uint32_t st = rd32(h, sb + 4);
	uint32_t fr = (st & 0xffffu) + 1u;
	uint32_t slots = st >> 16;
	int enq = 0;
	if (fr == slots) {
		enq = 1;
		fr = 0;
	}
	wr32(h, sb + 4, fr | (slots << 16));
	if (enq) {
		int c = rm_sbid_by_size(h->log2, size);
		if (rd64(h, H_SB(h, c)) == sb)
			wr64(h, H_SB(h, c), 0);
		q_enqueue(h, sb);
	}
}

int64_t rm_replay(rm_heap *h, const int64_t *ops, uint64_t n, uint64_t *out,
		  const uint32_t *pat)
{
	int64_t nulls = 0;
	for (uint64_t i = 0; i < n; i++) {
		if (ops[i] >= 0) {
			int st;
			uint64_t p = rm_malloc(h, (uint64_t)ops[i], &st);
			if (st == RM_HANG)
				return -1;
			out[i] = p;
			if (!p)
				nulls++;
			else if (pat)
				wr32(h, p, pat[i]);
		} else {
			uint64_t j = (uint64_t)(-ops[i] - 1);
			out[i] = 0;
			rm_free(h, out[j]);
		}
	}
	return nulls;
}

/* ---------------- deterministic trace values ---------------- */
uint64_t rm_mix(uint64_t seed, uint64_t gid, uint64_t round, uint64_t k)
{
	/* splitmix64 finaliser over a linear combination of the inputs */
	uint64_t z = seed * 0x9E3779B97F4A7C15ull + gid * 0xBF58476D1CE4E5B9ull +
		     round * 0x94D049BB133111EBull + k * 0xD6E8FEB86659FD93ull;
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

/* ================= multi-threaded restatement (CPU baseline) =============
 * Same lock-free protocol as kma.cl:150-373 / clIndexedQueue.cl:119-209,
 * with GCC __atomic builtins in place of atom_* + mem_fence.  Pointers are
 * real host addresses (base == img).
 */
#define A32(p) ((volatile uint32_t *)(uintptr_t)(p))
#define A64(p) ((volatile uint64_t *)(uintptr_t)(p))

static inline uint32_t ld32(uint64_t a) { return __atomic_load_n(A32(a), __ATOMIC_ACQUIRE); }
static inline uint64_t ld64(uint64_t a) { return __atomic_load_n(A64(a), __ATOMIC_ACQUIRE); }
static inline uint32_t cas32(uint64_t a, uint32_t cmp, uint32_t val)
{
	__atomic_compare_exchange_n(A32(a), &cmp, val, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
	return cmp;
}
static inline uint64_t cas64(uint64_t a, uint64_t cmp, uint64_t val)
{
	__atomic_compare_exchange_n(A64(a), &cmp, val, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
	return cmp;
}

typedef struct mt_heap {
	uint64_t base, qbase;
	uint32_t log2, sb, hdr, stride;
} mt_heap;

static inline uint64_t mt_idx2ptr(const mt_heap *m, uint32_t v)
{
	uint32_t i = (v >> 1) & 0xfffffu;
	return i ? m->qbase + ((uint64_t)(i - 1) << m->stride) : 0;
}
static inline uint32_t mt_ptr2idx(const mt_heap *m, uint64_t a)
{
	return (uint32_t)(((a - m->qbase) >> m->stride) + 1) << 1;
}

static void mt_enqueue(const mt_heap *m, uint64_t item)
{
	uint32_t tag = q_tag(ld32(item));
	__atomic_store_n(A32(item), q_ptr(0, tag - 1u), __ATOMIC_RELEASE);
	uint32_t idx = mt_ptr2idx(m, item);
	uint32_t tailidx;
	for (;;) {
		tailidx = ld32(m->base + 28);
		uint64_t tail = mt_idx2ptr(m, tailidx);
		uint32_t nextidx = ld32(tail);
		if (ld32(m->base + 28) == tailidx) {
			if (q_idx(nextidx) == 0) {
				if (cas32(tail, nextidx, q_ptr(idx, q_tag(nextidx))) == nextidx)
					break;
			} else {
				cas32(m->base + 28, tailidx, q_ptr(nextidx, q_tag(tailidx)));
			}
		}
	}
	cas32(m->base + 28, tailidx, q_ptr(idx, q_tag(tailidx)));
}

static uint64_t mt_dequeue(const mt_heap *m)
{
	for (;;) {
		uint32_t headidx = ld32(m->base + 24);
		uint64_t head = mt_idx2ptr(m, headidx);
		uint32_t tailidx = ld32(m->base + 28);
		uint32_t nextidx = ld32(head);
		if (headidx == ld32(m->base + 24)) {
			if (q_idx(headidx) == q_idx(tailidx)) {
				if (q_idx(nextidx) == 0)
					return 0;
				cas32(m->base + 28, tailidx, q_ptr(nextidx, q_tag(tailidx)));
			} else if (cas32(m->base + 24, headidx, q_ptr(nextidx, q_tag(headidx))) == headidx) {
				return head;
			}
		}
	}
}

static uint64_t mt_malloc(const mt_heap *m, uint64_t size)
{
	int block = rm_sbid_by_size(m->log2, size);
	if (block < 0 || block >= (int)(m->log2 - 1))
		return 0;
	uint64_t hsb = m->base + 32 + 8 * (uint64_t)block;
	uint64_t cursor;
	uint32_t slot;
	for (;;) {
		cursor = cas64(hsb, 0, POISON);
		if (cursor == 0) {
			cursor = mt_dequeue(m);
			if (!cursor) {
				__atomic_exchange_n(A64(hsb), 0, __ATOMIC_ACQ_REL);
				return 0;
			}
			uint32_t bsize = rm_size_by_sbid(m->log2, block);
			uint32_t slots = (uint32_t)rm_slots_by_size(m->log2, bsize);
			__atomic_store_n(A32(cursor + 8), bsize, __ATOMIC_RELAXED);
			for (uint32_t i = 0; i < slots; i += 32)
				__atomic_store_n(A32(cursor + m->sb - 4u * (i / 32u + 1u)), 0, __ATOMIC_RELAXED);
			__atomic_store_n(A32(cursor + 4), (slots << 16) | (slots - 1u), __ATOMIC_RELEASE);
			cas64(hsb, POISON, cursor);
			slot = slots - 1u;
			break;
		}
		if (cursor > POISON) {
			uint32_t so = ld32(cursor + 4);
			uint32_t fr = so & 0xffffu;
			if (fr == 0)
				continue;
			fr--;
			slot = fr;
			if (cas32(cursor + 4, so, fr | (so & 0xffff0000u)) != so)
				continue;
			if (fr == 0)
				__atomic_exchange_n(A64(hsb), 0, __ATOMIC_ACQ_REL);
			break;
		}
	}
	uint32_t slots = ld32(cursor + 4) >> 16;
	uint32_t bsize = ld32(cursor + 8);
	uint32_t start = slot;
	for (;;) {
		for (uint32_t b = slot; b < slots; b++) {
			uint64_t wa = cursor + m->sb - 4u * ((b >> 5) + 1u);
			uint32_t bit = 1u << (b & 31u);
			if (!(ld32(wa) & bit) &&
			    !(__atomic_fetch_or(A32(wa), bit, __ATOMIC_ACQ_REL) & bit)) {
				uint64_t p = cursor + 12u + (uint64_t)b * bsize;
				__atomic_store_n(A32(p), start, __ATOMIC_RELAXED);
				return p;
			}
		}
		slot = 0;
	}
}

static void mt_free(const mt_heap *m, uint64_t addr)
{
	if (!addr)
		return;
	uint64_t first_sb = m->base + m->hdr;
	uint64_t sb = first_sb + ((addr - first_sb) & ~(uint64_t)(m->sb - 1u));
	uint32_t size = ld32(sb + 8);
	uint64_t block = (addr - sb - 12u) / size;
	uint64_t wa = sb + m->sb - 4u * ((block >> 5) + 1u);
	uint32_t old = __atomic_fetch_and(A32(wa), ~(1u << (block & 31u)), __ATOMIC_ACQ_REL);
	__atomic_store_n(A32(addr), old, __ATOMIC_RELAXED);
	uint32_t so, ns;
	int enq;
	do {
		so = ld32(sb + 4);
		uint32_t fr = (so & 0xffffu) + 1u, slots = so >> 16;
		enq = fr == slots;
		ns = (enq ? 0u : fr) | (slots << 16);
	} while (cas32(sb + 4, so, ns) != so);
	if (enq) {
		int c = rm_sbid_by_size(m->log2, size);
		cas64(m->base + 32 + 8 * (uint64_t)c, sb, 0);
		mt_enqueue(m, sb);
	}
}

typedef struct mt_arg {
	const mt_heap *m;
	uint64_t wi0, wi1;
	uint32_t rounds, round0;
	int mode;
	uint32_t fixed;
	uint64_t seed;
	uint64_t ops, nulls, bad;
} mt_arg;

static void *mt_worker(void *p)
{
	mt_arg *a = (mt_arg *)p;
	for (uint64_t gid = a->wi0; gid < a->wi1; gid++) {
		for (uint32_t r = a->round0; r < a->round0 + a->rounds; r++) {
			uint64_t size;
			uint32_t hold;
			if (a->mode == 0) {
				size = a->fixed;
				hold = 0;
			} else {
				size = 1ull << (3 + rm_mix(a->seed, gid, r, 0) % 9);
				hold = (uint32_t)(rm_mix(a->seed, gid, r, 1) % 65);
			}
			uint64_t p = mt_malloc(a->m, size);
			if (!p) {
				a->nulls++;
				a->ops += 1;
				continue;
			}
			volatile uint32_t *w = (volatile uint32_t *)(uintptr_t)p;
			uint32_t pat = (uint32_t)(gid << 8) ^ r;
			if (a->mode == 0) {
				w[0] = (uint32_t)gid;
				w[1] = r;
				if (w[0] != (uint32_t)gid || w[1] != r)
					a->bad++;
			} else {
				uint32_t words = (uint32_t)(size / 4);
				w[0] = pat;
				for (uint32_t k = 0; k < hold; k++)
					w[1 + k % (words - 1)] = pat + k + 1;
				if (w[0] != pat)
					a->bad++;
			}
			mt_free(a->m, p);
			a->ops += 2;
		}
	}
	return NULL;
}

int rmt_run(uint32_t log2, uint32_t sblocks, uint32_t nthreads, uint64_t nwi,
	    uint32_t rounds, uint32_t round0, int mode, uint32_t fixed,
	    uint64_t seed, rmt_result *res)
{
	rm_heap h;
	uint64_t bytes = rm_heap_bytes(log2, sblocks);
	uint8_t *img = (uint8_t *)aligned_alloc(4096, (bytes + 4095) & ~4095ull);
	if (!img)
		return RM_EINVAL;
	memset(img, 0, bytes);
	if (rm_init(&h, log2, sblocks, (uint64_t)(uintptr_t)img, img) != RM_OK) {
		free(img);
		return RM_EINVAL;
	}
	mt_heap m = {h.base, h.base + h.hdr, log2, h.sb, h.hdr, log2};
	if (nthreads < 1)
		nthreads = 1;
	pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
	mt_arg *args = (mt_arg *)calloc(nthreads, sizeof(mt_arg));
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (uint32_t t = 0; t < nthreads; t++) {
		args[t] = (mt_arg){&m, nwi * t / nthreads, nwi * (t + 1) / nthreads,
				   rounds, round0, mode, fixed, seed, 0, 0, 0};
		pthread_create(&th[t], NULL, mt_worker, &args[t]);
	}
	memset(res, 0, sizeof(*res));
	for (uint32_t t = 0; t < nthreads; t++) {
		pthread_join(th[t], NULL);
		res->ops += args[t].ops;
		res->nulls += args[t].nulls;
		res->bad += args[t].bad;
	}
	clock_gettime(CLOCK_MONOTONIC, &t1);
	res->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
	free(th);
	free(args);
	free(img);
	return RM_OK;
}
