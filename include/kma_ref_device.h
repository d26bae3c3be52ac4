/*
 * kma_ref_device.h - device-side kma_malloc / kma_free on the reference heap
 * layout (HIP, gfx950).  Header-only so that user kernels can call it.
 *
 * Drop-in for the reference's device API
 *     void __global *malloc(__global struct clheap *heap, size_t);   clheap.h:38
 *     void free(__global struct clheap *heap, uintptr_t);            clheap.h:39
 * operating on a heap whose bytes are laid out exactly as the reference's
 * 64-bit `struct clheap` + superblocks (kma.h:71-85, clIndexedQueue.h:38-51):
 *   @0  u64 bytes      @8 u64 free.base   @16 u64 free.stride
 *   @24 u32 free.head  @28 u32 free.tail  @32 u64 sb[log2-1]
 *   superblock i at hdr + i*SB: u32 q.next | u32 state(slots<<16|free) |
 *   u32 size | blocks... | bitmap words growing down from the superblock end.
 * The superblock size is taken from free.stride, so one code path serves the
 * reference's 4 KiB build and 16 KiB heaps.
 *
 * The protocol is the reference's (kma.cl:150-373, clIndexedQueue.cl:119-209),
 * with agent-scope acquire/release atomics in place of atom_* + mem_fence and
 * four deliberate differences, none of which changes a serialized result:
 *  1. free() stores the old bitmap word into the block *before* it clears the
 *     bit; the reference stores it after (kma.cl:338), so another work-item
 *     may already own the block and have its first word overwritten.
 *  2. a refill clears the bitmap before it publishes the new state word
 *     (the reference publishes first, kma.cl:181-188, leaving a window in
 *     which a claimed bit can be wiped).
 *  3. a malloc that finds a 1-slot class attached with no free slot takes a
 *     fresh superblock instead of spinning (kma.cl:203-204 spins; on a
 *     64-wide wavefront the spinner can be the lane that holds the block).
 *  4. sizes 1..3 map to class 0 and size 0 returns NULL (the reference leaves
 *     the class uninitialised, kma.cl:129-135).
 *  5. a malloc whose class has no superblock attached while the free queue
 *     is empty returns NULL after plain loads instead of taking the class
 *     head's POISON lock to find the queue empty (same image, but without it
 *     every failing malloc of an exhausted heap queues on one word).
 *
 * Every wait is bounded.  Calls that may fail (reserve, the dequeue of a
 * refill) give up at a per-call watchdog or at a deadline the host arms in
 * the heap's stats trailer, and return NULL; calls that hold a reservation
 * or a released superblock (claim, free, enqueue) are lock-free and never
 * give up on a consistent heap.  A claim that still finds no free bit
 * returns its reserved slot.  Each give-up is counted in the trailer (256
 * bytes after the image, not part of the reference's layout).
 */
#ifndef KMA_REF_DEVICE_H
#define KMA_REF_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

struct clheap; /* reference-layout heap header; accessed by offsets */

namespace kma_ref {

static constexpr uint64_t POISON = 1u;
/* Every wait in this file gives up after this many ticks of the constant
 * 100 MHz wall clock; reaching it means the heap is corrupt or starved. */
#ifndef KMA_REF_WAIT_TICKS
#define KMA_REF_WAIT_TICKS 400000000ull /* 4 s */
#endif
static constexpr uint64_t WAIT_TICKS = KMA_REF_WAIT_TICKS;

/* Stats trailer: 256 bytes right after the heap image (at the header's
 * `bytes`, rounded up to 64), outside the reference's layout, so images stay
 * byte-exact.  u64 words: */
enum {
	T_DEADLINE = 0, /* wall-clock tick after which waits give up (0: none) */
	T_RESERVE = 1,  /* reserve gave up on its watchdog: malloc -> NULL */
	T_DLHIT = 2,    /* reserve gave up at the deadline: malloc -> NULL */
	T_CLAIM = 3,    /* claim found no free bit; its reserved slot was returned */
	T_DEQUEUE = 4,  /* dequeue gave up: the refill saw an empty queue */
	T_ENQUEUE = 5,  /* enqueue gave up (corrupt queue): a superblock is lost */
	T_NWORDS = 8
};

/* ---- agent-scope atomics ---- */
template <typename T>
__device__ __forceinline__ T ld(const T *p)
{
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st(T *p, T v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T cas(T *p, T cmp, T val)
{
	__hip_atomic_compare_exchange_strong(p, &cmp, val, __ATOMIC_ACQ_REL,
					     __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
	return cmp;
}
template <typename T>
__device__ __forceinline__ T xchg(T *p, T v)
{
	return __hip_atomic_exchange(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

/* Exponential back-off between retries of a spin or a failed CAS on a
 * shared word, so that thousands of spinning lanes do not starve the lane
 * they wait for of the word's atomic unit. */
__device__ __forceinline__ void backoff(uint32_t &k)
{
	for (uint32_t i = 0; i < k; i++)
		__builtin_amdgcn_s_sleep(4);
	k = k < 64u ? 2u * k : 64u;
}

/* ---- layout ---- */
struct view {
	char *h;       /* heap header */
	char *sb0;     /* first superblock == free.base */
	uint64_t *trl; /* stats trailer */
	uint32_t log2;
	uint32_t sbsz;
	uint32_t buckets;
};

__device__ __forceinline__ view make_view(clheap *heap)
{
	view v;
	v.h = reinterpret_cast<char *>(heap);
	v.log2 = (uint32_t)*reinterpret_cast<const uint64_t *>(v.h + 16);
	v.sbsz = 1u << v.log2;
	v.buckets = v.log2 - 1u;
	v.sb0 = v.h + 32 + 8u * v.buckets;
	uint64_t bytes = *reinterpret_cast<const uint64_t *>(v.h);
	v.trl = reinterpret_cast<uint64_t *>(v.h + ((bytes + 63u) & ~63ull));
	return v;
}

__device__ __forceinline__ void count(const view &v, int k)
{
	__hip_atomic_fetch_add(v.trl + k, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* true once the host-armed deadline has passed */
__device__ __forceinline__ bool past_deadline(const view &v)
{
	uint64_t dl = __hip_atomic_load(v.trl + T_DEADLINE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return dl != 0 && wall_clock64() > dl;
}

/* Wait bound of one call: the per-call watchdog and, where the caller may
 * give up (use_deadline), the heap's deadline; both read every 256 rounds.
 * expired(): 0 keep waiting, 1 watchdog, 2 deadline. */
struct waiter {
	uint64_t t0 = 0;
	uint32_t n = 0;
	__device__ __forceinline__ int expired(const view &v, bool use_deadline)
	{
		if ((++n & 255u) != 0)
			return 0;
		if (use_deadline && past_deadline(v))
			return 2;
		uint64_t t = wall_clock64();
		if (t0 == 0) {
			t0 = t;
			return 0;
		}
		return t - t0 > WAIT_TICKS ? 1 : 0;
	}
};

__device__ __forceinline__ uint32_t *q_head(const view &v) { return reinterpret_cast<uint32_t *>(v.h + 24); }
__device__ __forceinline__ uint32_t *q_tail(const view &v) { return reinterpret_cast<uint32_t *>(v.h + 28); }
__device__ __forceinline__ uint64_t *cls_head(const view &v, int c) { return reinterpret_cast<uint64_t *>(v.h + 32 + 8 * c); }
__device__ __forceinline__ uint32_t *sb_next(char *sb) { return reinterpret_cast<uint32_t *>(sb); }
__device__ __forceinline__ uint32_t *sb_state(char *sb) { return reinterpret_cast<uint32_t *>(sb + 4); }
__device__ __forceinline__ uint32_t *sb_size(char *sb) { return reinterpret_cast<uint32_t *>(sb + 8); }
__device__ __forceinline__ uint32_t *sb_word(const view &v, char *sb, uint32_t b)
{
	return reinterpret_cast<uint32_t *>(sb + v.sbsz) - ((b >> 5) + 1u);
}

/* Tagged 32-bit queue word: bit0 poison, bits1..16 index, bits22..32 tag. */
__device__ __forceinline__ uint32_t q_ptr(uint32_t i, uint32_t t) { return (i & 0x1fffeu) | (((t + 1u) & 0x7ffu) << 22); }
__device__ __forceinline__ uint32_t q_tag(uint32_t w) { return (w >> 22) & 0x7ffu; }
__device__ __forceinline__ uint32_t q_idx(uint32_t w) { return (w & 0x1fffeu) >> 1; }

__device__ __forceinline__ char *idx2ptr(const view &v, uint32_t w)
{
	uint32_t i = (w >> 1) & 0xfffffu;
	return i ? v.sb0 + ((uint64_t)(i - 1u) << v.log2) : nullptr;
}
__device__ __forceinline__ uint32_t ptr2idx(const view &v, char *p)
{
	return (uint32_t)((((uint64_t)(p - v.sb0)) >> v.log2) + 1u) << 1;
}

/* ---- size classes (kma.cl:55-148) ---- */
__device__ __forceinline__ uint32_t size_by_sbid(uint32_t log2, int sbid)
{
	int buckets = (int)log2 - 1;
	if (sbid > buckets || sbid < 0)
		return 0;
	int e = sbid + 2;
	if (e <= (int)(log2 >> 1))
		return 1u << e;
	uint32_t items = (1u << log2) >> e;
	uint32_t ab = (items >> 3) + ((items & 7u) ? 1u : 0u);
	ab = (ab + 3u) & ~3u;
	return (((1u << log2) - 12u - ab) / items) & ~3u;
}

__device__ __forceinline__ uint32_t slots_by_size(uint32_t log2, uint32_t size)
{
	uint32_t space = (1u << log2) - 12u;
	uint32_t s = space / size;
	if (s & 0x1fu)
		s += 32u;
	s = (s >> 3) & ~3u;
	return (space - s) / size;
}

__device__ __forceinline__ int sbid_by_size(uint32_t log2, uint64_t size)
{
	if (size == 0 || size > (1ull << log2) - 16u)
		return -1;
	if (size < 4)
		return 0;
	/* highest set bit of size>>2 below 2^buckets, then walk to the class
	 * whose block size is the smallest one >= size */
	uint32_t ss = (uint32_t)(size >> 2) & ((1u << (log2 - 1u)) - 1u);
	int sbid = ss ? 31 - __clz(ss) : 0;
	for (;;) {
		uint32_t t = size_by_sbid(log2, sbid);
		if (size > t && t > 0)
			sbid++;
		else if (size < size_by_sbid(log2, sbid - 1))
			sbid--;
		else
			return sbid;
	}
}

/* ---- indexed queue (clIndexedQueue.cl:119-209) ---- */
/* Enqueue is lock-free (a failed CAS means another operation went
 * through), so it has no deadline: a superblock is never dropped from a
 * consistent queue.  Only a corrupt queue (a cycle) can keep it looping;
 * then the watchdog ends the call and the loss is counted. */
__device__ inline void q_enqueue(const view &v, char *item)
{
	uint32_t tag = q_tag(ld(sb_next(item)));
	st(sb_next(item), q_ptr(0, tag - 1u));
	uint32_t idx = ptr2idx(v, item);
	uint32_t tailw;
	waiter wt;
	uint32_t bo = 1;
	for (;;) {
		if (wt.expired(v, false)) {
			count(v, T_ENQUEUE);
			return;
		}
		tailw = ld(q_tail(v));
		char *tail = idx2ptr(v, tailw);
		uint32_t nextw = ld(sb_next(tail));
		if (ld(q_tail(v)) != tailw)
			continue;
		if (q_idx(nextw) == 0) {
			if (cas(sb_next(tail), nextw, q_ptr(idx, q_tag(nextw))) == nextw)
				break;
			backoff(bo);
		} else {
			cas(q_tail(v), tailw, q_ptr(nextw, q_tag(tailw)));
		}
	}
	cas(q_tail(v), tailw, q_ptr(idx, q_tag(tailw)));
}

/* Dequeue; NULL when the queue is empty, or when the wait for a consistent
 * snapshot ends (watchdog or deadline; counted): nothing was taken then. */
__device__ inline char *q_dequeue(const view &v)
{
	waiter wt;
	uint32_t bo = 1;
	for (;;) {
		if (wt.expired(v, true)) {
			count(v, T_DEQUEUE);
			return nullptr;
		}
		uint32_t headw = ld(q_head(v));
		char *head = idx2ptr(v, headw);
		uint32_t tailw = ld(q_tail(v));
		uint32_t nextw = ld(sb_next(head));
		if (ld(q_head(v)) != headw)
			continue;
		if (q_idx(headw) == q_idx(tailw)) {
			if (q_idx(nextw) == 0)
				return nullptr;
			cas(q_tail(v), tailw, q_ptr(nextw, q_tag(tailw)));
		} else if (cas(q_head(v), headw, q_ptr(nextw, q_tag(headw))) == headw) {
			return head;
		} else {
			backoff(bo);
		}
	}
}

/* The queue holds no superblock to pop: the dummy node is both head and
 * tail and has no successor (q_dequeue's empty case, by plain loads). */
__device__ __forceinline__ bool q_empty(const view &v)
{
	uint32_t headw = ld(q_head(v));
	uint32_t tailw = ld(q_tail(v));
	if (q_idx(headw) != q_idx(tailw))
		return false;
	return q_idx(ld(sb_next(idx2ptr(v, headw)))) == 0;
}

/* ---- reserve + claim (kma.cl:150-273) ---- */
/* A malloc that finds its class with no superblock attached and the free
 * queue empty returns NULL after plain loads: the reference takes the class
 * head's POISON lock for that (CAS 0 -> POISON, empty dequeue, release),
 * which leaves the same image but serializes every failing malloc of an
 * exhausted heap on one word. */
__device__ inline char *reserve(const view &v, int block, uint32_t *slot)
{
	uint64_t *hp = cls_head(v, block);
	if (past_deadline(v)) {
		count(v, T_DLHIT);
		return nullptr;
	}
	waiter wt;
	uint32_t bo = 1;
	for (;;) {
		int e = wt.expired(v, true);
		if (e) {
			count(v, e == 2 ? T_DLHIT : T_RESERVE);
			return nullptr;
		}
		uint64_t cur = ld(hp);
		if (cur == 0) {
			if (q_empty(v))
				return nullptr;
			cur = cas(hp, (uint64_t)0, POISON);
		}
		if (cur == 0) {
			char *sb = q_dequeue(v);
			if (!sb) {
				xchg(hp, (uint64_t)0);
				return nullptr;
			}
			uint32_t bsize = size_by_sbid(v.log2, block);
			uint32_t slots = slots_by_size(v.log2, bsize);
			st(sb_size(sb), bsize);
			for (uint32_t i = 0; i < slots; i += 32)
				st(sb_word(v, sb, i), 0u);
			st(sb_state(sb), (slots << 16) | (slots - 1u));
			cas(hp, POISON, (uint64_t)(uintptr_t)sb);
			*slot = slots - 1u;
			return sb;
		}
		if (cur > POISON) {
			char *sb = reinterpret_cast<char *>((uintptr_t)cur);
			uint32_t so = ld(sb_state(sb));
			uint32_t fr = so & 0xffffu;
			if (fr == 0) {
				/* 1-slot class attached full: detach and refill
				 * rather than wait (see note 3 above). */
				if ((so >> 16) == 1u)
					cas(hp, cur, (uint64_t)0);
				backoff(bo);
				continue;
			}
			fr--;
			if (cas(sb_state(sb), so, (so & 0xffff0000u) | fr) != so) {
				backoff(bo);
				continue;
			}
			if (fr == 0)
				xchg(hp, (uint64_t)0);
			*slot = fr;
			return sb;
		}
		backoff(bo);
	}
}

/* One slot of `sb` back into its free count (kma.cl:340-371): when every
 * slot is free the superblock is reset, unlinked from its class head and
 * enqueued.  Lock-free: a failed CAS means another count update went in. */
__device__ inline void release_slot(const view &v, char *sb)
{
	uint32_t so, ns, bo = 1;
	bool enq;
	for (;;) {
		so = ld(sb_state(sb));
		uint32_t fr = (so & 0xffffu) + 1u, slots = so >> 16;
		enq = fr == slots;
		ns = (enq ? 0u : fr) | (slots << 16);
		if (cas(sb_state(sb), so, ns) == so)
			break;
		backoff(bo);
	}
	if (enq) {
		int c = sbid_by_size(v.log2, ld(sb_size(sb)));
		cas(cls_head(v, c), (uint64_t)(uintptr_t)sb, (uint64_t)0);
		q_enqueue(v, sb);
	}
}

/* Claim the first free bit from `slot` upward, wrapping to 0.  The
 * reservation guarantees a free bit in a consistent heap; if none turns up
 * before the watchdog, the reserved slot is returned and the claim counted
 * (NULL), so the superblock can still empty and go back to the queue. */
__device__ inline char *claim(const view &v, char *sb, uint32_t slot)
{
	uint32_t slots = ld(sb_state(sb)) >> 16;
	uint32_t bsize = ld(sb_size(sb));
	uint32_t start = slot;
	waiter wt;
	do {
		for (uint32_t b = slot; b < slots; b++) {
			uint32_t *w = sb_word(v, sb, b);
			uint32_t bit = 1u << (b & 31u);
			if (ld(w) & bit)
				continue;
			if (!(__hip_atomic_fetch_or(w, bit, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) & bit)) {
				char *p = sb + 12 + (uint64_t)b * bsize;
				*reinterpret_cast<volatile uint32_t *>(p) = start;
				return p;
			}
		}
		slot = 0;
	} while (!wt.expired(v, false));
	count(v, T_CLAIM);
	release_slot(v, sb);
	return nullptr;
}

} /* namespace kma_ref */

/* Allocate `size` bytes from a reference-layout heap; NULL when the size has
 * no class (size 0 or > SB-16) or no superblock is left (kma.cl:279-303). */
__device__ inline void *kma_malloc(clheap *heap, size_t size)
{
	using namespace kma_ref;
	view v = make_view(heap);
	int block = sbid_by_size(v.log2, size);
	if (block < 0 || block >= (int)v.buckets)
		return nullptr;
	uint32_t slot;
	char *sb = reserve(v, block, &slot);
	if (!sb)
		return nullptr;
	return claim(v, sb, slot);
}

/* Return a block to a reference-layout heap; NULL is a no-op
 * (kma.cl:305-373). */
__device__ inline void kma_free(clheap *heap, void *ptr)
{
	using namespace kma_ref;
	if (!ptr)
		return;
	view v = make_view(heap);
	char *p = reinterpret_cast<char *>(ptr);
	uint64_t off = (uint64_t)(p - v.sb0);
	char *sb = v.sb0 + (off & ~(uint64_t)(v.sbsz - 1u));
	uint32_t size = ld(sb_size(sb));
	uint32_t b = (uint32_t)((uint64_t)(p - sb - 12) / size);
	uint32_t *w = sb_word(v, sb, b);
	/* old word into the block first, then release the bit */
	*reinterpret_cast<volatile uint32_t *>(p) = ld(w);
	__hip_atomic_fetch_and(w, ~(1u << (b & 31u)), __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
	release_slot(v, sb);
}

#endif /* KMA_REF_DEVICE_H */
